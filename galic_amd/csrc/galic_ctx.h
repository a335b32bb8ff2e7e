// galic_ctx.h -- device-resident state behind the C ABI.
//
// The reference keeps this state in globals (allvars.h:308-407, 844-868) and in
// per-rank copies that MPI_Allreduce sums (main.c:259-285).  On one GPU every
// grid lives in HBM once.  Layouts:
//   particles      SoA arrays of length N (pos/vel split per component)
//   Loc grids      [type][field][cell], field = m, r, t, q, p (orbit-averaged)
//                  and [type][field][cell], field = r, t, p, q (instantaneous)
//   stacks         [type][which][Nstack], quadtree levels as STACKOFFSET
//   orbit slots    [orbit][cell][field], one private response per orbit (cell-major)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "galic_device.h"

namespace galic {

// stack indices in Ctx::stacks
enum StackId {
  ST_MT = 0,   // DGs_MassTarget
  ST_EMT,      // EGs_MassTarget
  ST_ER,       // EGs_EgyTarget_r
  ST_ET,       // EGs_EgyTarget_t
  ST_EP,       // EGs_EgyTarget_p
  ST_EQ,       // EGs_EgyTarget_q
  ST_RM,       // DGs_MassResponse
  ST_RSR,      // EGs_EgyResponseRS
  ST_RST,      // EGs_EgyResponseTS
  ST_RSQ,      // EGs_EgyResponseQS
  ST_RSP,      // EGs_EgyResponsePS
  ST_EMR,      // EGs_MassResponse
  ST_IR,       // EGs_EgyResponse_r
  ST_IT,       // EGs_EgyResponse_t
  ST_IP,       // EGs_EgyResponse_p
  ST_IQ,       // EGs_EgyResponse_q
  ST_DIST,     // DGs_Distance (type-independent, stored per type for simplicity)
  ST_COUNT
};

// per-node weights and residuals used by the incremental merit
enum WeightId {
  W_A = 0,  // min(dist, R200/20) on cut-A nodes (density cut), else 0
  W_SR,     // Tm/Er on cut-A nodes with Er > 0, else 0 (then W_ST, W_SQ, W_SP)
  W_ST,
  W_SQ,
  W_SP,
  W_B,      // 1 on cut-B nodes (cut on EGs_MassResponse), else 0
  W_COUNT
};

// orbit-averaged fields (slot order = reference argument order m, r, t, q, p)
enum { F_M = 0, F_R, F_T, F_Q, F_P };
// instantaneous fields (add_to_energy_grid order r, t, p, q)
enum { I_R = 0, I_T, I_P, I_Q };

struct Particles {
  int n = 0;
  double *x = nullptr, *y = nullptr, *z = nullptr;
  double *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *mass = nullptr, *tint = nullptr, *vesc = nullptr;
  double *vr2 = nullptr, *vt2 = nullptr, *vp2 = nullptr, *vq2 = nullptr;
  int *type = nullptr, *orbits = nullptr;
  double *disp = nullptr;  // target vr2, vt2, vp2, vq2 per particle (set_particles.c:158-161)
};

// The per-launch state of the FP32 path's orbit integration: the stream and events it runs
// on, the log arena, the per-orbit inputs and outputs, the response logs' tables and counters,
// and the sort buffers of a group of trials.  A context holds two sets, the active one in its
// own fields and the other in Ctx::alt: one group's orbits integrate on one set's stream while
// the previous group's reps are judged with the other's (swap_orbit_set in galic_api.hip).
#define GALIC_ORBIT_SET_FIELDS(X) \
  X(hipStream_t, stream) X(hipEvent_t, ev0) X(hipEvent_t, ev1) X(void *, slots) X(size_t, slot_bytes) \
  X(int, max_slot_orbits) X(int, max_orbits) X(int *, o_pidx) X(int *, o_type) X(double *, o_x) X(double *, o_y) \
  X(double *, o_z) X(double *, o_vx) X(double *, o_vy) X(double *, o_vz) X(double *, o_mass) X(double *, o_tint) \
  X(double *, o_tall) X(double *, o_e0) X(double *, o_e1) X(int *, o_norbit) X(int *, o_status) \
  X(long long *, o_steps) X(int4 *, o_bbox) X(unsigned int *, counter) X(unsigned long long *, orbit_stats) \
  X(int, max_sort_trials) X(unsigned int *, t_key) X(unsigned int *, t_key2) X(int *, t_ord) X(int *, t_ord2) \
  X(int *, t_inv) X(void *, sort_tmp) X(size_t, sort_tmp_bytes) X(int *, log_blocks) X(size_t, log_blocks_len) \
  X(int *, log_n) X(float *, log_tot) X(int, log_n_len) X(int, log_M) X(unsigned int *, log_next) X(int *, log_q) \
  X(float4 *, log_v)
struct OrbitSet {
#define GALIC_OS_FIELD(T, f) T f{};
  GALIC_ORBIT_SET_FIELDS(GALIC_OS_FIELD)
#undef GALIC_OS_FIELD
};

struct Ctx {
  galic_params p;
  int device = 0;
  hipStream_t stream = nullptr;
  DevModel M;
  int FG_Ngrid = 0, Ngrid = 0, Nstack = 0, L = 0, Nbin = 0;
  bool targets_set[GALIC_NTYPES] = {false, false, false, false};
  bool cuts_ready = false;

  // tables
  double *FG_Pot = nullptr, *FG_DPotDR = nullptr, *FG_DPotDz = nullptr;
  double *FG_Vstream = nullptr;  // [type][FG_Ngrid]
  double *DG_CellSize = nullptr;
  float2 *FG_Acc_f = nullptr;
  float *FG_Vstream_f = nullptr;
  unsigned int vstream_nz = 0;     // bit t: type t's FP32 streaming table has a nonzero node
  float *DG_Cell1_f = nullptr;     // [Nbin] cell size by max(iR, iz), where the table has that form
  bool cell1_ok = false;
  float *DG_CellSize_f = nullptr;

  // grids
  double *stacks = nullptr;     // [type][ST_COUNT][Nstack]
  double *weights = nullptr;    // [type][W_COUNT][Nstack]
  unsigned char *tcut = nullptr;  // [type][16x16 tile, Morton index] levels of the tile's cut-A nodes
  double *locA = nullptr;       // orbit-averaged Loc [type][5][Ngrid]
  double *locA_d = nullptr;     // deltas
  double *locI = nullptr;       // instantaneous Loc [type][4][Ngrid]
  double *locI_d = nullptr;
  double *locM = nullptr;       // EG_MassLoc [type][Ngrid]
  long long *locA_q = nullptr;  // fixed-point forms used by the FP32 path
  long long *locA_dq = nullptr;
  long long *locI_q = nullptr;
  long long *locI_dq = nullptr;
  long long *locM_q = nullptr;
  double *qinv = nullptr;       // 1/scale of the 5 orbit-averaged then 4 instantaneous fields
  int *tovs_dev = nullptr;      // velocity structure per type, on the device
  double *t_vk = nullptr;       // K trial velocities per particle, SoA [3][max_ktrials]
  double *t_dFk = nullptr;      // merit change of every trial of the last FP32 call [max_ktrials]
  double *t_vinst = nullptr;    // instantaneous merit change of every trial (sparse tile merit) [max_ktrials]
  size_t max_ktrials = 0;
  int last_ntrial = 0, last_k = 0;  // trials and K of the last FP32 call
  double *fit_part = nullptr;   // partial cut sums of the global fit [type-1][kFitBlocks][5]
  double *fit = nullptr;        // [type][16] S, Sdisp_r,t,p,q, Srelfac, Srelfac_count, Sall, raw sums..., firstrun
  Particles P;

  // orbit slots (FP64 path: double; FP32 path: float)
  void *slots = nullptr;
  size_t slot_bytes = 0;
  int max_slot_orbits = 0;

  // per-orbit scratch
  int max_orbits = 0;
  int *o_pidx = nullptr, *o_type = nullptr;
  double *o_x = nullptr, *o_y = nullptr, *o_z = nullptr;
  double *o_vx = nullptr, *o_vy = nullptr, *o_vz = nullptr;
  double *o_mass = nullptr, *o_tint = nullptr;
  double *o_tall = nullptr, *o_e0 = nullptr, *o_e1 = nullptr;
  int *o_norbit = nullptr, *o_status = nullptr;
  long long *o_steps = nullptr;
  int4 *o_bbox = nullptr;
  unsigned int *counter = nullptr;
  unsigned long long *orbit_stats = nullptr;  // [0] steps of the launch, [1] first failed orbit + 1, [2] its status,
                                              // [3] orbits whose response log overflowed

  // per-trial scratch
  int max_trials = 0;
  int *t_idx = nullptr;
  double *t_v = nullptr;  // [3][max_trials]
  int *cutB_of = nullptr; // [type][Ngrid]
  int *t_accept = nullptr;
  double *t_dF = nullptr;
  int *t_best = nullptr;
  unsigned int *t_key = nullptr, *t_key2 = nullptr;  // trial sort keys (expected orbit steps)
  int *t_ord = nullptr, *t_ord2 = nullptr;           // trial order, longest expected orbit first
  int *t_inv = nullptr;                              // chunk position of each trial
  int max_sort_trials = 0;                           // capacity of the five arrays above
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int *psteps = nullptr;  // per particle: steps of its current orbit (FP32 path's length prediction)
  double2 *merit_bw = nullptr;  // FP32 merit: (response - target, node weight) [type][5][Nstack]
  float2 *merit_bwf = nullptr;  // the same in FP32 (the 16x16-tile merit)
  // response logs of the batched FP32 path (arena carved from `slots`)
  int *log_blocks = nullptr;     // [orbit][log_M]
  size_t log_blocks_len = 0;
  int *log_n = nullptr;          // [orbit]
  float *log_tot = nullptr;      // [orbit][5] sums of each orbit's record values
  // region commit of logs into the Loc grids (galic_commit.hip)
  unsigned int *cm_cnt = nullptr;    // [2][type][region] counts, cursors
  unsigned int *cm_start = nullptr;  // [type][region] bucket starts
  int cm_nk = 0;                     // types x regions of the above
  int4 *cm_chunks = nullptr;         // (type * regions + region, lo, hi, 0)
  size_t cm_chunks_len = 0;
  uint2 *cm_buf = nullptr;           // bucket entries
  size_t cm_buf_len = 0;
  // reps of a group captured in one hipGraph (the layout above made on the device)
  unsigned int *cm_aux = nullptr;          // [0] chunks laid out, [1] room exceeded
  unsigned long long *cm_bound = nullptr;  // [rep of the group] entries a commit may need
  int cm_bound_len = 0;
  int use_graph = 1;                       // capture a group's reps (merit, choice, commit, fit)
  long long acc_graph_launches = 0;
  int log_n_len = 0;
  int log_M = 64;                // block-table entries per orbit (doubled when an orbit needs more)
  unsigned int *log_next = nullptr;
  int *log_q = nullptr;          // the last launch's arena: quads, then values
  float4 *log_v = nullptr;
  double log_est_blocks = 12;    // running estimate of arena blocks per orbit
  float log_scale[5] = {0, 0, 0, 0, 0};  // record value -> exact int64 sum units (a power of two)
  double log_unit[5] = {0, 0, 0, 0, 0};   // its inverse
  double log_to_loc[5] = {0, 0, 0, 0, 0}; // int64 sum units -> Loc fixed point
  double max_mass = 0, max_vesc = 0;  // of the particles, bound the log's fixed point

  // target dispersion tables of the sampled targets, per type: FG_Disp_r [FG_Nbin], then FG_DispZ,
  // FG_DispPhi, FG_tilted_vR2_prime, FG_tilted_vz2_prime [FG_Nbin^2 each] (disp_fields.c)
  double *disp_tab[GALIC_NTYPES] = {nullptr, nullptr, nullptr, nullptr};
  double disp_beta[GALIC_NTYPES] = {0, 0, 0, 0};   // HaloBetaParameter / BulgeBetaParameter
  double disp_bound[GALIC_NTYPES] = {0, 0, 0, 0};  // bound of every dispersion the tables give
  float last_target_ms = 0;                          // kernel time of the last target sampling
  double *cs_scratch = nullptr;  // FP64 merit: coarse pyramid levels in HBM for grids above level 7
  size_t cs_scratch_len = 0;
  long long *g32_scratch = nullptr;  // FP32 merit: two int64 grids per workgroup for grids above level 7
  size_t g32_scratch_len = 0;
  unsigned int *tile_list = nullptr;  // FP32 merit, tile form: the records' tile lists, per workgroup
  size_t tile_list_len = 0;
  unsigned int *tile_aux = nullptr;
  double *tsrc = nullptr;        // sources of the table fill (galic_set_tree_sources): [4][tsrc_n] x, y, z, mass
  int tsrc_n = 0;
  double tsrc_lo[3] = {0, 0, 0}, tsrc_hi[3] = {0, 0, 0};  // their extent
  double *lookup_buf = nullptr;  // galic_field_at: staging of positions, accelerations, potentials
  size_t lookup_len = 0;         // points it holds   // [0] largest records of a particle's logs, [1] list overflow flag
  unsigned long long *merit_prof = nullptr;  // [16] phase times of the sparse tile merit (galic_fp32_merit_profile)
  int merit_form = 0;                 // 0 automatic, 1 32x32 tiles at every level, 2 LDS grids / HBM scratch (no tiles), 3 sparse 32x32 tiles, 4 16x16 tiles
  OrbitSet alt;                       // the other orbit set (GALIC_ORBIT_SET_FIELDS); allocated on first use
  bool alt_ready = false;
  int set_parity = 0;                 // 1 while the sets are swapped
  int overlap = -1;                   // FP32 reps: integrate the next group's orbits while judging a group
                                      // (-1 when a call needs more than one launch, 0 never, 1 always)
  int merit_wg_per_cu = 0;            // FP32 merit kernel: workgroups per CU at most (0: as many as fit)
  int max_group_trials = 0;           // FP32 reps: at most this many trials per orbit launch (0: the arena's)

  // timing of the last launches (ms), measured with HIP events on ctx stream
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_orbit_ms = 0, last_eval_ms = 0, last_fit_ms = 0;
  double acc_orbit_ms = 0, acc_eval_ms = 0, acc_fit_ms = 0;
  long long acc_log_relaunches = 0;
  long long acc_orbit_launches = 0;
  long long acc_orbits = 0, acc_steps = 0;
  int fp32_variant = 0;   // orbit kernel variant (one so far)
  int fp32_bpc = 4;       // persistent workgroups per CU for k_orbit32 (autotuned)
  int fp32_bpc_overlap = 0;  // the same for the launches beside a group's judging (0: fp32_bpc)
  int fp32_nfresh = 0;    // device trials per particle drawn fresh (optimize()'s distribution)
  bool fp32_tuned = false;
  double fp32_scale[5];  // fixed-point scales per field (orbit-averaged)
  double fp32_iscale;    // fixed-point scale of the instantaneous fields
};

}  // namespace galic
