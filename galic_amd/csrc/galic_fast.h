// galic_fast.h -- argument blocks of the FP32 batched path (galic_fast.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "galic_ctx.h"
#include "galic_kernels.h"

namespace galic {

// FP32 forms of the model constants (reciprocals and log2 are fine here).
struct FastModel {
  int use_table, add_halo, add_bulge;
  int tovs[GALIC_NTYPES];
  int fg_nbin, dg_nbin;
  float fg_invRmin, fg_invlog2Fac, dg_invRmin, dg_invlog2Fac;
  float halo_GM, halo_a, halo_eps, offgrid_fac;
  float bulge_GM, bulge_a, bulge_eps;
  float fo_v200, fc;
  long long max_steps;
  const float2 *acc;  // (DPotDR, DPotDz) per node
  const float *vstream;  // [type][node] streaming velocity
  unsigned int vstream_nz;  // bit t: type t's streaming table has a nonzero node (else it is not read)
  const float *cellsize;
  const float *cell1;      // or null: cellsize[iz * nb + iR] = cell1[max(iR, iz)] (staged in LDS)
  double qscale[kFields];  // fixed-point scale per orbit-averaged field
};

constexpr int kCell1Max = 1024;  // density-grid cells per side the LDS cell-size table holds

struct OrbitArgs32 {
  int n;
  unsigned int *counter;
  const double *ox, *oy, *oz, *ovx, *ovy, *ovz, *omass, *otint;
  const int *otype;
  float *slots;        // private FP32 response per orbit, or null
  size_t slot_stride;  // floats per slot (5*Ngrid)
  int slot_base;
  long long *locq;     // fixed-point Loc grids [type][5][Ngrid] for direct atomics, or null
  double *tall;
  int *norbit, *status;
  long long *steps;
  int4 *bbox;
  unsigned long long *stats;
  float *fin;  // final x, y, z, vx, vy, vz per orbit, or null
  // response logs (k_orbit32_log): every flushed 2x2 run is appended as a record
  // (quad index, 4 cells x 5 fields of fac * weight in FP32) to the orbit's blocks
  int *log_blocks;       // [orbit][log_M] arena block of each log block
  int *log_n;            // [orbit] records written (-1: the log overflowed)
  float *log_tot;        // [orbit][5] sum of the orbit's record values per field
  int log_M;             // block-table entries per orbit
  int *log_q;            // arena: [block][LOG_B] quad index
  float4 *log_v;         // arena: [block][LOG_B][5] values (a record's five fields side by side)
  unsigned int *log_next;  // next free arena block
  unsigned int log_cap;    // arena blocks
};

// records per log block
constexpr int LOG_B = 64;
// arena blocks a lane claims at once (the blocks a lane has not used when the
// launch ends are left empty: at most LOG_CHUNK - 1 per lane)
constexpr int LOG_CHUNK = 8;
// Where two grids fit in LDS, the merit kernel's grids hold row iz of the
// response grid at iz * (nb + MERIT_ROW_PAD): a record's cells one row apart
// then fall in different LDS banks.
constexpr int MERIT_ROW_PAD = 2;
// Threads of a merit workgroup: 512; 1024 for 128^2 grids, where one workgroup
// fills the LDS of a CU and brings all the waves the CU runs, and for 256^2 and
// 512^2 grids, where more waves share one particle's long chain of blocks.  The per-wave merit
// sums have room for MERIT_MAX_WAVES waves.
constexpr int MERIT_MAX_WAVES = 16;
__host__ __device__ constexpr int merit_threads(int level) { return level >= 7 ? 1024 : 512; }

struct EvalArgs32 {
  Particles P;
  int K;
  int trial_base;       // block b evaluates trial trial_base + b (b = block_base + blockIdx.x)
  int block_base;
  const int *slot_pos;  // ... whose orbits sit at chunk position slot_pos[b] (null: b)
  long long *gscratch;  // grids above 128^2: two int64 grids per workgroup of the launch, zero between uses
  int nblk;             // tile form: trials of the launch (a workgroup takes trials b, b + gridDim.x, ...)
  unsigned int *tlist;  // tile form: per workgroup, tlist_cap entries of the records' tile lists
  int tlist_cap;
  unsigned int *tile_err;  // tile form: set when a particle's lists exceed tlist_cap
  uint2 *tlist2;           // 16x16-tile form: per workgroup, tlist_cap entries (record index, log)
  const float *log_tot;    // [orbit][5] sums of each log's values (k_orbit32_log)
  int no_deltas;           // the taken trial's deltas are left to the region commit (k_commit_*)
  const unsigned char *tcut;  // 16x16-tile form: [type][tile] levels of the tile's cut-A nodes (bit u: L-u
                              // inside the tile, bit 5: a cut-A node above the tile)
  const int *work_order;   // 16x16-tile form: trials in the order workgroups take them
  int *work_next;          // ... from this counter (zero at launch)
  unsigned long long *prof;  // sparse tile form, optional: time per phase and sizes, summed over particles
  int tbl_lds;               // sparse tile form: the block tables staged in LDS (else read from log_blocks)
  const double *vinst;       // sparse tile form: instantaneous merit change of each trial (k_trial_inst)
  const int *tidx;
  const double *tvx, *tvy, *tvz;  // trial velocities [i*K + k]
  const double *stacks, *fit;
  const double2 *merit_bw;  // (response - target, weight) per node [type][5][Nstack]
  const float2 *merit_bwf;  // the same in FP32 (the 16x16-tile form)
  const int *cutB_of;
  const int *log_blocks, *log_n;  // the chunk's response logs (OrbitArgs32)
  int log_M;
  const int *log_q;
  const float4 *log_v;
  float log_scale[kFields];   // record value -> exact int64 sum units (a power of two)
  double log_unit[kFields];   // int64 sum units -> response units (its inverse)
  double log_to_loc[kFields]; // int64 sum units -> Loc fixed point
  const int4 *bbox;
  const int *norbit;
  const long long *osteps;  // steps of each orbit of the chunk
  int *psteps;              // per particle: steps of its current orbit (updated)
  long long *locA_dq, *locI_dq;
  int *accept;
  double *dF;
  double *dFk;  // merit change of trial k of trial ti at ti * K + k
  double MType[GALIC_NTYPES];
  double fac[GALIC_NTYPES][4];
  double lowdisp;
  double qscale[kFields];
  double iscale;
};

__global__ void k_orbit32(FastModel F, OrbitArgs32 A);
__global__ void k_orbit32_log(FastModel F, OrbitArgs32 A);
__global__ void k_draw_trials(DevModel M, int ntrial, int K, int nfresh, const int *tidx, Particles P,
                              const double *disp, double sigma_frac, double maxvel, uint32_t seed, uint32_t stream,
                              const int *tovs_of_type, double *vtx, double *vty, double *vtz);
__global__ void k_prepare_ktrials(int ntrial, int K, const int *order, const int *tidx, const double *vtx,
                                  const double *vty, const double *vtz, Particles P, OrbitIn O);
__global__ void k_store_steps(int n, const long long *osteps, int *psteps);
__global__ void k_invert_order(int n, const int *ord, int *inv);
__global__ void k_energy32(DevModel M, int n, const double *ox, const double *oy, const double *oz, const double *ovx,
                           const double *ovy, const double *ovz, const float *fin, double *e0, double *e1);
int sort_trials_by_steps(int ntrial, const int *tidx, const int *psteps, unsigned int *key, unsigned int *key2,
                         int *ord, int *ord2, void **tmp, size_t *tmp_bytes, hipStream_t stream);
int sort_trials_by_records(int n, int K, const int *slot_pos, const int *log_n, unsigned int *key, unsigned int *key2,
                           int *ord, int *ord2, void **tmp, size_t *tmp_bytes, hipStream_t stream);
__global__ void k_pack_merit(const double *stacks, const double *weights, double2 *bw, float2 *bwf, int Nstack);
template <bool USE_O, bool GMEM, int LF>
__global__ void k_eval_log32(DevModel M, EvalArgs32 A);
// Tile form of the FP32 merit: the response grid is cut into tiles of TILE x TILE cells,
// each log's records are listed by the tiles their cells fall in, and the tiles the logs
// touch are summed in LDS one at a time.  Rows of a tile grid are padded to TILE_RS cells.
constexpr int TILE = 32;
constexpr int TILE_RS = TILE + 2;
constexpr int TILE_CELLS = TILE * TILE_RS;
constexpr int TILE_MAX_K = 63;  // list entries hold the log index in 8 bits
// Nodes of one tile's pyramid below its root: 32 x 32 cells, then 16 x 16, 8 x 8, 4 x 4 and
// the 2 x 2 block nodes (levels L .. L-4).
constexpr int TILE_NODES = 1024 + 256 + 64 + 16 + 4;
// LDS of k_eval_tile32 beyond its static arrays: per-wave partials [K][16] double, the
// per-trial pyramid above the tiles [K][Nstack(L-5)] and block sums [2][K][4] int64, K + 2 tile
// grids (two for the current orbit, alternating between tiles), the logs touching each tile
// [ntile] uint64, tile counts, offsets and the active list [3 * ntile + 1] int, the block
// tables [K+1][log_M] int, and with BWL the tile's (residual, weight) pairs [TILE_NODES] double2.
__host__ __device__ inline size_t tile_lds_bytes(int L, int K, int log_M, bool bwl) {
  const size_t nst = GALIC_NSTACK(L - 5), ntile = (size_t)1 << (2 * (L - 5));
  return (size_t)MERIT_MAX_WAVES * K * 8 + (size_t)K * nst * 8 + 8 * (size_t)K * 8 +
         (size_t)(K + 2) * TILE_CELLS * 8 + ntile * 8 + (3 * ntile + 1) * 4 + (size_t)(K + 1) * log_M * 4 +
         (bwl ? (size_t)TILE_NODES * 16 + 16 : 0);
}
template <int LF, bool BWL>
__global__ void k_eval_tile32(DevModel M, EvalArgs32 A);
// LDS of k_eval_sparse32 beyond its static arrays: per-wave partials [K][16] double, the
// per-trial pyramid above the tiles [K][Nstack(L-5)] and the 4x4-cell node sums [2][K][64]
// int64, K + 2 tile grids, the logs touching each tile [ntile] uint64, tile counts, offsets
// and the active list [3 * ntile + 1] int, the block tables [K+1][log_M] int (when they are
// staged in LDS: a table grows with the longest orbit's log), the marked
// nodes [2][8] uint32, the tile's pairs of cells and 2x2 nodes [1280] and of the coarser
// nodes [2][84] double2.
__host__ __device__ inline size_t sparse_lds_bytes(int L, int K, int log_M, bool tbl_lds) {
  const size_t nst = GALIC_NSTACK(L - 5), ntile = (size_t)1 << (2 * (L - 5));
  return (size_t)MERIT_MAX_WAVES * K * 8 + (size_t)K * nst * 8 + 128 * (size_t)K * 8 +
         (size_t)(K + 2) * TILE_CELLS * 8 + ntile * 8 + (3 * ntile + 1) * 4 +
         (tbl_lds ? (size_t)(K + 1) * log_M * 4 : 0) + 64 + 16 + (size_t)(1280 + 168) * 16;
}
template <int LF>
__global__ void k_eval_sparse32(DevModel M, EvalArgs32 A);
// 16x16-tile form of the FP32 merit (k_eval_t16): int32 tile grids in the particle's own fixed
// point, one wave per field, all trials of a tile per lane; at most T16_MAX_K trials.
constexpr int T16 = 16;
constexpr int T16_MAX_K = 16;
constexpr int T16_MAX_LEVEL = 10;  // Morton codes of up to 64 x 64 tiles; open ancestors of up to 8 levels
// int32 words at the front of its LDS: the K+1 grids (the current orbit's, then the trials')
// of a tile's five fields (the tile counts share them while the lists are built).
__host__ __device__ inline int t16_grid_words(int K, int L) {
  const int a = (K + 1) * kFields * T16 * T16, b = 1 << (2 * (L - 4));
  return a > b ? a : b;
}
// words of the per-tile log masks: two tiles a word for up to 15 trials (KR the trial capacity)
__host__ __device__ inline int t16_mask_words(int ntile, int KR) { return KR <= 15 ? (ntile + 1) / 2 : ntile; }
// LDS of k_eval_t16 beyond its static arrays: the grids, the list offsets [ntile + 1], the
// active tiles [ntile] (16-bit) and the masks.
__host__ __device__ inline size_t t16_lds_bytes(int L, int K, int KR) {
  const int ntile = 1 << (2 * (L - 4));
  return 4 * ((size_t)t16_grid_words(K, L) + ntile + 1 + (ntile + 1) / 2 + t16_mask_words(ntile, KR));
}
template <int LF, int KR>
__global__ void k_eval_t16(DevModel M, EvalArgs32 A);
__global__ void k_trial_inst(DevModel M, EvalArgs32 A, double *vinst);
__global__ void k_apply_best(DevModel M, EvalArgs32 A);
__global__ void k_log_totals(int n, int K, const int *slot_pos, const int *log_n, unsigned int *maxsum);

// Logs into the fixed-point Loc grids through region buckets (galic_commit.hip).
struct CommitArgs {
  int n;                    // items of the launch: trials (with accept) or orbits (without)
  int K, trial_base;        // with accept: trial b is trial_base + b, its orbits at slot_pos[b] * (K + 1)
  const int *slot_pos;
  const int *accept;        // taken trial + 1 per trial; null: item b is orbit b, added whole
  const int *tidx, *ptype;  // with accept: particle of each trial, type of each particle
  const int *otype;         // without: type of each orbit (indexed by b)
  int orbit_base;           // without: item b is orbit orbit_base + b of the logs
  const int *log_blocks, *log_n;
  int log_M;
  const int *log_q;
  const float4 *log_v;
  int L, gsh;               // grid level; log2 of a region's side (min(L, 5))
  unsigned int *gcnt;       // [type][region] entries (k_commit_count adds)
  unsigned int *gcur;       // [type][region] cursors (zero before k_commit_scatter)
  const unsigned int *gstart;  // [type][region] first entry of the bucket
  uint2 *gbuf;              // entries (arena record index, sign: 0 add, 1 take away)
  double q[kFields];        // Loc fixed-point units per response unit
  const unsigned int *gerr; // set: the layout did not fit its room, k_commit_scatter writes nothing
};
constexpr int kCommitRegionLog2 = 5;
__global__ void k_commit_count(CommitArgs A);
__global__ void k_commit_scatter(CommitArgs A);
__global__ void k_commit_regions(CommitArgs A, const int4 *chunks, long long *dst, const unsigned int *nchunk);
__global__ void k_commit_layout(const unsigned int *gcnt, int nk, int chunk, unsigned int *gstart, int4 *chunks,
                                unsigned int *nchunk, unsigned int cap_entries, unsigned int cap_chunks,
                                unsigned int *err);
__global__ void k_commit_bound(int n, int K, int first, int rep_size, const int *slot_pos, const int *log_n,
                               unsigned long long *rep_bound);
__global__ void k_fill_u32(unsigned int *p, unsigned int v, size_t n);
// Trials [0, n) of a merit launch in order of decreasing record count (ties by index), one
// workgroup of 1024, n <= kOrderMax
constexpr int kOrderMax = 16384;
__global__ void k_record_order(int n, int K, const int *slot_pos, const int *log_n, int *ord);

__global__ void k_commit_q(long long *locq, long long *dq, double *loc, size_t n_per_field, int nfield,
                           const double *inv_scale, size_t n);
__global__ void k_inst_init_q(DevModel M, Particles P, long long *locIq, long long *locMq, int ngrid, double iscale,
                              double mscale);

}  // namespace galic
