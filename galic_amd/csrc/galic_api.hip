// galic_api.hip -- host side of the C ABI (include/galic_hip.h): context
// management, uploads, and the driver sequences of the reference optimiser
// (main.c:93-102, 151-246, 808-874) expressed as launches on one HIP stream.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <string>
#include <vector>

#include "../../include/galic_hip.h"
#include "galic_ctx.h"
#include "galic_fast.h"
#include "galic_targets.h"
#include "galic_kernels.h"
#include "galic_handle.h"

namespace {

thread_local std::string g_err;
galic_ctx *g_bound = nullptr;

int set_err(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                                \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if(e_ != hipSuccess)                                                                            \
      return set_err(e_ == hipErrorOutOfMemory ? GALIC_ERR_NOMEM : GALIC_ERR_HIP, "%s: %s (%s:%d)", \
                     #expr, hipGetErrorString(e_), __FILE__, __LINE__);                             \
  } while(0)

// HIP status as a library status (for paths that must release memory before returning)
int hip_ok(hipError_t e) {
  if(e == hipSuccess) return GALIC_OK;
  return set_err(e == hipErrorOutOfMemory ? GALIC_ERR_NOMEM : GALIC_ERR_HIP, "%s", hipGetErrorString(e));
}

#define CHK(expr)          \
  do {                     \
    int r_ = (expr);       \
    if(r_ != GALIC_OK)     \
      return r_;           \
  } while(0)

template <typename T>
int dalloc(T **ptr, size_t count) {
  *ptr = nullptr;
  if(count == 0) return GALIC_OK;
  HIPCHK(hipMalloc((void **)ptr, count * sizeof(T)));
  HIPCHK(hipMemset(*ptr, 0, count * sizeof(T)));
  return GALIC_OK;
}

template <typename T>
void dfree(T *&ptr) {
  if(ptr) (void)hipFree((void *)ptr);
  ptr = nullptr;
}

inline int nblocks(size_t n, int bs) { return (int)((n + bs - 1) / bs); }

// Newton solve for the spacing factor (grid.c:353-374 and the FG/EG copies).
double solve_fac(int nbin, double rin, double rmax, double *rmin) {
  int iter = 0;
  double dfac, fac = 1.02, y = rmax / rin;
  do {
    double f = log((pow(fac, nbin) - 1.0) / (y * (fac - 1.0)));
    double df = nbin / (fac - 1.0 / pow(fac, nbin - 1)) - 1.0 / (fac - 1.0);
    dfac = -f / df;
    fac += dfac;
    iter++;
    if(iter > 1000) return -1;
  } while(fabs(dfac) > 1.0e-8 * fac);
  *rmin = rin / (fac - 1.0);
  return fac;
}

void smooth_stack_host(double *data, int maxlevel) {
  for(int l = maxlevel - 1; l >= 0; l--) {
    int n = 1 << l;
    for(int i = 0; i < n; i++)
      for(int j = 0; j < n; j++)
        data[GALIC_STACKOFFSET(l, i, j)] =
            data[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j)] + data[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j)] +
            data[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j + 1)] + data[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j + 1)];
  }
}

// The set of nodes where calc_stack_difference's recursion stops
// (grid.c:1090: l >= maxlevel or ref1 < thresh).  Positions never move, so
// the reference's cuts are the same at every call.
void mark_cut(const double *ref, double thresh, int l, int i, int j, int L, std::vector<char> &cut) {
  int o = GALIC_STACKOFFSET(l, i, j);
  if(l >= L || ref[o] < thresh) {
    cut[o] = 1;
    return;
  }
  for(int ii = 2 * i; ii <= 2 * i + 1; ii++)
    for(int jj = 2 * j; jj <= 2 * j + 1; jj++) mark_cut(ref, thresh, l + 1, ii, jj, L, cut);
}

bool type_present(const galic_params &p, int t) { return p.NType[t] > 0; }

int alloc_orbit_scratch(galic::Ctx &c, int n) {
  if(n <= c.max_orbits) return GALIC_OK;
  dfree(c.o_pidx); dfree(c.o_type); dfree(c.o_x); dfree(c.o_y); dfree(c.o_z);
  dfree(c.o_vx); dfree(c.o_vy); dfree(c.o_vz); dfree(c.o_mass); dfree(c.o_tint);
  dfree(c.o_tall); dfree(c.o_e0); dfree(c.o_e1); dfree(c.o_norbit); dfree(c.o_status);
  dfree(c.o_steps); dfree(c.o_bbox);
  CHK(dalloc(&c.o_pidx, n)); CHK(dalloc(&c.o_type, n));
  CHK(dalloc(&c.o_x, n)); CHK(dalloc(&c.o_y, n)); CHK(dalloc(&c.o_z, n));
  CHK(dalloc(&c.o_vx, n)); CHK(dalloc(&c.o_vy, n)); CHK(dalloc(&c.o_vz, n));
  CHK(dalloc(&c.o_mass, n)); CHK(dalloc(&c.o_tint, n));
  CHK(dalloc(&c.o_tall, n)); CHK(dalloc(&c.o_e0, n)); CHK(dalloc(&c.o_e1, n));
  CHK(dalloc(&c.o_norbit, n)); CHK(dalloc(&c.o_status, n));
  CHK(dalloc(&c.o_steps, n)); CHK(dalloc(&c.o_bbox, n));
  c.max_orbits = n;
  return GALIC_OK;
}

// The sort buffers of a group of trials (part of the active orbit set).
int alloc_sort_scratch(galic::Ctx &c, int n) {
  if(n <= c.max_sort_trials) return GALIC_OK;
  dfree(c.t_key); dfree(c.t_key2); dfree(c.t_ord); dfree(c.t_ord2); dfree(c.t_inv);
  c.max_sort_trials = 0;
  CHK(dalloc(&c.t_inv, n));
  CHK(dalloc(&c.t_key, n));
  CHK(dalloc(&c.t_key2, n));
  CHK(dalloc(&c.t_ord, n));
  CHK(dalloc(&c.t_ord2, n));
  c.max_sort_trials = n;
  return GALIC_OK;
}

int alloc_trial_scratch(galic::Ctx &c, int n) {
  CHK(alloc_sort_scratch(c, n));
  if(n <= c.max_trials) return GALIC_OK;
  dfree(c.t_idx); dfree(c.t_v); dfree(c.t_accept); dfree(c.t_dF); dfree(c.t_best);
  CHK(dalloc(&c.t_idx, n));
  CHK(dalloc(&c.t_v, 3 * (size_t)n));
  CHK(dalloc(&c.t_accept, n));
  CHK(dalloc(&c.t_dF, n));
  CHK(dalloc(&c.t_best, n));
  c.max_trials = n;
  return GALIC_OK;
}

// Slots hold one private response (5 grids) per orbit of a trial chunk.
int alloc_slots(galic::Ctx &c, int norbits, size_t bytes_per_orbit) {
  size_t need = (size_t)norbits * bytes_per_orbit;
  if(need <= c.slot_bytes) return GALIC_OK;
  if(c.slots) (void)hipFree(c.slots);
  c.slots = nullptr;
  c.slot_bytes = 0;
  HIPCHK(hipMalloc(&c.slots, need));
  HIPCHK(hipMemset(c.slots, 0, need));
  c.slot_bytes = need;
  return GALIC_OK;
}

// Exchanges the active orbit set with the other one (see galic::OrbitSet).
void swap_orbit_set(galic::Ctx &c) {
#define GALIC_OS_SWAP(T, f) std::swap(c.f, c.alt.f);
  GALIC_ORBIT_SET_FIELDS(GALIC_OS_SWAP)
#undef GALIC_OS_SWAP
  c.set_parity ^= 1;
}

// The other set's stream, events and counters (its arrays grow on demand when it is active).
int ensure_alt_set(galic::Ctx &c) {
  if(c.alt_ready) return GALIC_OK;
  swap_orbit_set(c);
  int rc = GALIC_OK;
  c.log_M = 64;
  if(hipStreamCreate(&c.stream) != hipSuccess || hipEventCreate(&c.ev0) != hipSuccess ||
     hipEventCreate(&c.ev1) != hipSuccess)
    rc = set_err(GALIC_ERR_HIP, "creating the second orbit stream failed");
  if(rc == GALIC_OK) rc = dalloc(&c.counter, 1);
  if(rc == GALIC_OK) rc = dalloc(&c.orbit_stats, 4);
  if(rc == GALIC_OK) rc = dalloc(&c.log_next, 1);
  swap_orbit_set(c);
  if(rc == GALIC_OK) c.alt_ready = true;
  return rc;
}

// Frees the active orbit set's memory, events and stream.
void free_orbit_set(galic::Ctx &c) {
  if(c.stream) (void)hipStreamSynchronize(c.stream);
  dfree(c.o_pidx); dfree(c.o_type); dfree(c.o_x); dfree(c.o_y); dfree(c.o_z);
  dfree(c.o_vx); dfree(c.o_vy); dfree(c.o_vz); dfree(c.o_mass); dfree(c.o_tint);
  dfree(c.o_tall); dfree(c.o_e0); dfree(c.o_e1); dfree(c.o_norbit); dfree(c.o_status);
  dfree(c.o_steps); dfree(c.o_bbox); dfree(c.counter); dfree(c.orbit_stats);
  dfree(c.t_key); dfree(c.t_key2); dfree(c.t_ord); dfree(c.t_ord2); dfree(c.t_inv);
  dfree(c.log_blocks); dfree(c.log_n); dfree(c.log_tot); dfree(c.log_next);
  if(c.sort_tmp) (void)hipFree(c.sort_tmp);
  c.sort_tmp = nullptr;
  if(c.slots) (void)hipFree(c.slots);
  c.slots = nullptr;
  if(c.ev0) (void)hipEventDestroy(c.ev0);
  if(c.ev1) (void)hipEventDestroy(c.ev1);
  if(c.stream) (void)hipStreamDestroy(c.stream);
  c.ev0 = c.ev1 = nullptr;
  c.stream = nullptr;
}

galic::OrbitIn orbit_in(galic::Ctx &c) {
  galic::OrbitIn O;
  O.pidx = c.o_pidx;
  O.ox = c.o_x; O.oy = c.o_y; O.oz = c.o_z;
  O.ovx = c.o_vx; O.ovy = c.o_vy; O.ovz = c.o_vz;
  O.omass = c.o_mass; O.otint = c.o_tint; O.otype = c.o_type;
  return O;
}

galic::OrbitArgs orbit_args(galic::Ctx &c, int n) {
  galic::OrbitArgs A;
  memset(&A, 0, sizeof(A));
  A.n = n;
  A.counter = c.counter;
  A.ox = c.o_x; A.oy = c.o_y; A.oz = c.o_z;
  A.ovx = c.o_vx; A.ovy = c.o_vy; A.ovz = c.o_vz;
  A.omass = c.o_mass; A.otint = c.o_tint; A.otype = c.o_type;
  A.tall = c.o_tall; A.e0 = c.o_e0; A.e1 = c.o_e1;
  A.norbit = c.o_norbit; A.status = c.o_status; A.steps = c.o_steps; A.bbox = c.o_bbox;
  A.slot_stride = (size_t)galic::kFields * c.Ngrid;
  A.stats = c.orbit_stats;
  return A;
}

// Persistent grid: enough workgroups to fill every CU several times over, never
// more than the orbits need.
int orbit_grid(int n, int bs, int per_cu = 8) {
  int dev;
  (void)hipGetDevice(&dev);
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  int want = cus * per_cu;
  int need = nblocks(n, bs);
  return std::max(1, std::min(want, need));
}

// Steps and the first failure of an orbit launch (orbit_response.c:449-456 exits on the step cap).
int finish_orbit_launch(galic::Ctx &c, int n) {
  unsigned long long st[3];
  HIPCHK(hipMemcpyAsync(st, c.orbit_stats, sizeof(st), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  c.acc_steps += (long long)st[0];
  c.acc_orbits += n;
  if(st[1]) {
    long long status = (long long)st[2];
    int o = (int)(st[1] - 1);
    if(status == GALIC_ERR_STEPCAP)
      return set_err(GALIC_ERR_STEPCAP, "orbit %d exceeded the step cap of %lld steps", o, (long long)c.M.max_steps);
    return set_err(GALIC_ERR_ARG, "orbit %d has a non-positive or non-finite timestep", o);
  }
  return GALIC_OK;
}

int launch_orbit64(galic::Ctx &c, galic::OrbitArgs A) {
  if(A.n <= 0) return GALIC_OK;
  HIPCHK(hipMemsetAsync(c.counter, 0, sizeof(unsigned int), c.stream));
  HIPCHK(hipMemsetAsync(c.orbit_stats, 0, 4 * sizeof(unsigned long long), c.stream));
  HIPCHK(hipEventRecord(c.ev0, c.stream));
  const int bs = 256;
  hipLaunchKernelGGL(galic::k_orbit64, dim3(orbit_grid(A.n, bs)), dim3(bs), 0, c.stream, c.M, A);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c.ev1, c.stream));
  HIPCHK(hipEventSynchronize(c.ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  c.last_orbit_ms = ms;
  c.acc_orbit_ms += ms;
  c.acc_orbit_launches++;
  return finish_orbit_launch(c, A.n);
}

galic::FitArgs fit_args(galic_ctx *ctx) {
  galic::Ctx &c = ctx->c;
  galic::FitArgs F;
  F.L = c.L;
  for(int t = 0; t < GALIC_NTYPES; t++) {
    F.present[t] = (t > 0 && type_present(c.p, t)) ? 1 : 0;
    F.tovs[t] = c.p.TypeOfVelocityStructure[t];
  }
  F.stacks = c.stacks;
  F.weights = c.weights;
  F.locA = c.locA;
  F.locI = c.locI;
  F.locM = c.locM;
  F.fit = c.fit;
  F.lowdisp = c.p.LowerDispLimit;
  return F;
}

// Cut-A weights (density cut, fixed by DGs_MassTarget) and, once EGs_MassResponse
// is known, the cut-B flags and the fine cell -> cut-B node map.
int build_weights(galic_ctx *ctx, bool with_cutB) {
  galic::Ctx &c = ctx->c;
  const int Ns = c.Nstack, L = c.L, nb = c.Nbin;
  ctx->weights.assign((size_t)GALIC_NTYPES * galic::W_COUNT * Ns, 0.0);
  ctx->cutB_of.assign((size_t)GALIC_NTYPES * c.Ngrid, 0);
  // per 16 x 16-cell tile (Morton index), the levels of its cut-A nodes: bit u for level L-u
  // inside the tile (u = 0..4), bit 5 when the tile lies inside a cut-A node above its level
  const int lt = L - 4, ntile = lt >= 0 ? 1 << (2 * lt) : 0;
  std::vector<unsigned char> tcut((size_t)GALIC_NTYPES * ntile, 0);
  std::vector<double> emr(Ns);
  for(int t = 1; t < GALIC_NTYPES; t++) {
    if(!type_present(c.p, t) || !c.targets_set[t]) continue;
    double *W = &ctx->weights[(size_t)t * galic::W_COUNT * Ns];
    const std::vector<double> &Tm = ctx->targets[t][0];
    double thr_d = c.p.MinParticlesPerBinForDensityMeasurement * c.p.MType[t] / c.p.NType[t];
    std::vector<char> cutA(Ns, 0);
    mark_cut(Tm.data(), thr_d, 0, 0, 0, L, cutA);
    for(int tz = 0; tz < (1 << std::max(lt, 0)) && lt >= 0; tz++)
      for(int tr = 0; tr < (1 << lt); tr++) {
        unsigned char m = 0;
        for(int u = 0; u <= 4; u++) {
          const int n = 16 >> u, l = L - u;
          for(int i = 0; i < n && !(m & (1 << u)); i++)
            for(int j = 0; j < n; j++)
              if(cutA[GALIC_STACKOFFSET(l, tz * n + i, tr * n + j)]) {
                m |= (unsigned char)(1 << u);
                break;
              }
        }
        for(int l = 0; l < lt; l++)
          if(cutA[GALIC_STACKOFFSET(l, tz >> (lt - l), tr >> (lt - l))]) m |= 32;
        int code = 0;
        for(int b = 0; b < 6; b++) code |= (((tr >> b) & 1) << (2 * b)) | (((tz >> b) & 1) << (2 * b + 1));
        tcut[(size_t)t * ntile + code] = m;
      }
    const double rlim = c.p.R200 / 20.0;
    for(int o = 0; o < Ns; o++) {
      if(!cutA[o]) continue;
      double d = ctx->distance[o];
      W[galic::W_A * Ns + o] = (d < rlim) ? d : rlim;  // min() of grid.c:1078-1084
      // S-term weights Tm/E_x for E_x > 0 (grid.c:1128-1133), fields r, t, q, p
      const int e_of[4] = {2, 3, 5, 4};  // targets[] order: Tm, MT, Er, Et, Ep, Eq
      for(int k = 0; k < 4; k++) {
        double e = ctx->targets[t][e_of[k]][o];
        W[(galic::W_SR + k) * Ns + o] = (e > 0) ? Tm[o] / e : 0.0;
      }
    }
    if(with_cutB) {
      HIPCHK(hipMemcpy(emr.data(), c.stacks + ((size_t)t * galic::ST_COUNT + galic::ST_EMR) * Ns,
                       sizeof(double) * Ns, hipMemcpyDeviceToHost));
      double thr_s = c.p.MinParticlesPerBinForDispersionMeasurement * c.p.MType[t] / c.p.NType[t];
      std::vector<char> cutB(Ns, 0);
      mark_cut(emr.data(), thr_s, 0, 0, 0, L, cutB);
      for(int o = 0; o < Ns; o++) W[galic::W_B * Ns + o] = cutB[o] ? 1.0 : 0.0;
      int *cb = &ctx->cutB_of[(size_t)t * c.Ngrid];
      for(int iz = 0; iz < nb; iz++)
        for(int iR = 0; iR < nb; iR++) {
          int node = -1;
          for(int l = 0; l <= L && node < 0; l++) {
            int o = GALIC_STACKOFFSET(l, iz >> (L - l), iR >> (L - l));
            if(cutB[o]) node = o;
          }
          cb[iz * nb + iR] = node;
        }
    }
  }
  HIPCHK(hipMemcpy(c.weights, ctx->weights.data(), sizeof(double) * ctx->weights.size(), hipMemcpyHostToDevice));
  if(ntile > 0) {
    if(!c.tcut) CHK(dalloc(&c.tcut, tcut.size()));
    HIPCHK(hipMemcpy(c.tcut, tcut.data(), tcut.size(), hipMemcpyHostToDevice));
  }
  if(with_cutB)
    HIPCHK(hipMemcpy(c.cutB_of, ctx->cutB_of.data(), sizeof(int) * ctx->cutB_of.size(), hipMemcpyHostToDevice));
  return GALIC_OK;
}

// The fit's kernels on c.stream (the first call also sets the cuts, which waits for the stream).
int global_fit_launch(galic_ctx *ctx) {
  galic::Ctx &c = ctx->c;
  galic::FitArgs F = fit_args(ctx);
  if(c.L >= 5) {
    // grids of 32 x 32 cells and more: the stacks in tiles, then the levels above them
    const int ntile = (c.Nbin / 32) * (c.Nbin / 32);
    hipLaunchKernelGGL(galic::k_fit_tiles, dim3(ntile, 30), dim3(256), 0, c.stream, F);
    hipLaunchKernelGGL(galic::k_fit_top, dim3(30), dim3(256), 0, c.stream, F);
  } else {
    hipLaunchKernelGGL(galic::k_fit_stacks, dim3(30), dim3(256), 0, c.stream, F);
  }
  HIPCHK(hipGetLastError());
  if(!c.cuts_ready) {
    HIPCHK(hipStreamSynchronize(c.stream));
    CHK(build_weights(ctx, true));
    c.cuts_ready = true;
  }
  hipLaunchKernelGGL(galic::k_fit_partial, dim3(galic::kFitBlocks, 3), dim3(256), 0, c.stream, F, c.fit_part,
                     galic::kFitBlocks);
  hipLaunchKernelGGL(galic::k_fit_final, dim3(3), dim3(64), 0, c.stream, F, (const double *)c.fit_part,
                     galic::kFitBlocks);
  HIPCHK(hipGetLastError());
  return GALIC_OK;
}

int global_fit(galic_ctx *ctx) {
  galic::Ctx &c = ctx->c;
  HIPCHK(hipEventRecord(c.ev0, c.stream));
  CHK(global_fit_launch(ctx));
  HIPCHK(hipEventRecord(c.ev1, c.stream));
  HIPCHK(hipEventSynchronize(c.ev1));
  HIPCHK(hipEventElapsedTime(&c.last_fit_ms, c.ev0, c.ev1));
  c.acc_fit_ms += c.last_fit_ms;
  return GALIC_OK;
}

int commit(galic_ctx *ctx) {
  galic::Ctx &c = ctx->c;
  size_t na = (size_t)GALIC_NTYPES * galic::kFields * c.Ngrid, ni = (size_t)GALIC_NTYPES * 4 * c.Ngrid;
  hipLaunchKernelGGL(galic::k_commit, dim3(nblocks(na, 256)), dim3(256), 0, c.stream, c.locA, c.locA_d, na);
  hipLaunchKernelGGL(galic::k_commit, dim3(nblocks(ni, 256)), dim3(256), 0, c.stream, c.locI, c.locI_d, ni);
  HIPCHK(hipGetLastError());
  return global_fit(ctx);
}


// The FP32 cell sizes of the density grid, and their per-index table when every cell's size is
// the size of cell (m, m), m = max(iR, iz) (the geometry of grid.c:329-487 gives this: a cell's
// size is the larger of its R and z widths, both the same function of the index).
int upload_cell_sizes_f(galic::Ctx &c, const double *cs) {
  const int nb = c.Nbin;
  std::vector<float> f(c.Ngrid), d(nb);
  for(int i = 0; i < c.Ngrid; i++) f[i] = (float)cs[i];
  for(int m = 0; m < nb; m++) d[m] = f[(size_t)m * nb + m];
  bool ok = nb <= galic::kCell1Max;
  for(int k = 0; k < nb && ok; k++)
    for(int j = 0; j < nb && ok; j++) ok = f[(size_t)k * nb + j] == d[std::max(j, k)];
  HIPCHK(hipMemcpy(c.DG_CellSize_f, f.data(), sizeof(float) * c.Ngrid, hipMemcpyHostToDevice));
  if(ok) HIPCHK(hipMemcpy(c.DG_Cell1_f, d.data(), sizeof(float) * nb, hipMemcpyHostToDevice));
  c.cell1_ok = ok;
  return GALIC_OK;
}

galic::FastModel fast_model(galic::Ctx &c) {
  galic::FastModel F;
  memset(&F, 0, sizeof(F));
  const galic_params &p = c.p;
  F.use_table = c.M.use_table;
  F.add_halo = c.M.add_halo;
  F.add_bulge = c.M.add_bulge;
  for(int t = 0; t < GALIC_NTYPES; t++) F.tovs[t] = p.TypeOfVelocityStructure[t];
  F.fg_nbin = p.FG_Nbin;
  F.dg_nbin = c.Nbin;
  F.fg_invRmin = (float)(1.0 / p.FG_Rmin);
  F.fg_invlog2Fac = (float)(1.0 / log2(p.FG_Fac));
  F.dg_invRmin = (float)(1.0 / p.DG_Rmin);
  F.dg_invlog2Fac = (float)(1.0 / log2(p.DG_Fac));
  F.halo_GM = (float)(p.G * p.Halo_Mass);
  F.halo_a = (float)p.Halo_A;
  F.halo_eps = (float)(1.0e-6 * p.Halo_A);
  F.offgrid_fac = (float)c.M.halo_offgrid_fac;
  F.bulge_GM = (float)(p.G * p.Bulge_Mass);
  F.bulge_a = (float)p.Bulge_A;
  F.bulge_eps = (float)(1.0e-6 * p.Bulge_A);
  F.fo_v200 = (float)(p.TimeStepFactorOrbit * p.V200);
  F.fc = (float)p.TimeStepFactorCellCross;
  F.max_steps = c.M.max_steps;
  F.acc = c.FG_Acc_f;
  F.vstream = c.FG_Vstream_f;
  F.vstream_nz = c.vstream_nz;
  F.cellsize = c.DG_CellSize_f;
  F.cell1 = c.cell1_ok ? c.DG_Cell1_f : nullptr;
  for(int f = 0; f < 5; f++) F.qscale[f] = c.fp32_scale[f];
  return F;
}

galic::OrbitArgs32 orbit_args32(galic::Ctx &c, int n) {
  galic::OrbitArgs32 A;
  memset(&A, 0, sizeof(A));
  A.n = n;
  A.counter = c.counter;
  A.ox = c.o_x; A.oy = c.o_y; A.oz = c.o_z;
  A.ovx = c.o_vx; A.ovy = c.o_vy; A.ovz = c.o_vz;
  A.omass = c.o_mass; A.otint = c.o_tint; A.otype = c.o_type;
  A.tall = c.o_tall; A.norbit = c.o_norbit; A.status = c.o_status; A.steps = c.o_steps; A.bbox = c.o_bbox;
  A.slot_stride = (size_t)galic::kFields * c.Ngrid;
  A.stats = c.orbit_stats;
  return A;
}

int launch_orbit32(galic::Ctx &c, galic::OrbitArgs32 A) {
  if(A.n <= 0) return GALIC_OK;
  HIPCHK(hipMemsetAsync(c.counter, 0, sizeof(unsigned int), c.stream));
  HIPCHK(hipMemsetAsync(c.orbit_stats, 0, 4 * sizeof(unsigned long long), c.stream));
  HIPCHK(hipEventRecord(c.ev0, c.stream));
  const int bs = 256;
  hipLaunchKernelGGL(galic::k_orbit32, dim3(orbit_grid(A.n, bs, c.fp32_bpc)), dim3(bs), 0, c.stream, fast_model(c),
                     A);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c.ev1, c.stream));
  HIPCHK(hipEventSynchronize(c.ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  c.last_orbit_ms = ms;
  c.acc_orbit_ms += ms;
  c.acc_orbit_launches++;
  return finish_orbit_launch(c, A.n);
}

// Exact sums of the response logs: the merit kernel turns each FP32 record
// value into an int64 at log_scale[f] = 2^k units per response unit and adds
// integers, so D = new - old is exact and independent of the order.  k is the
// largest power that keeps an orbit's total of a field (the particle's mass,
// times at most vmax^2 for a moment) below 2^60.
void log_setup(galic::Ctx &c) {
  const double vmax = c.max_vesc > 0 ? 2 * c.max_vesc : 10 * c.p.V200;
  for(int f = 0; f < galic::kFields; f++) {
    const double bound = std::max(c.max_mass, 1e-300) * (f == 0 ? 1.0 : vmax * vmax);
    int e = 0;
    frexp(bound, &e);  // bound < 2^e
    const int k = std::max(-120, std::min(120, 60 - e));
    c.log_scale[f] = ldexpf(1.0f, k);
    c.log_unit[f] = ldexp(1.0, -k);
    c.log_to_loc[f] = c.fp32_scale[f] * ldexp(1.0, -k);
  }
}

size_t log_block_bytes() { return (size_t)galic::LOG_B * (sizeof(int) + galic::kFields * sizeof(float4)); }

// The log arena (the `slots` buffer): reallocated only when it is below half of
// what a call could use, so repeated calls reuse one allocation.
int ensure_arena(galic::Ctx &c, size_t want, size_t need) {
  if(c.slot_bytes >= need && 2 * c.slot_bytes >= want) return GALIC_OK;
  if(c.slots) (void)hipFree(c.slots);
  c.slots = nullptr;
  c.slot_bytes = 0;
  const size_t bytes = std::max(want, need);
  HIPCHK(hipMalloc(&c.slots, bytes));
  c.slot_bytes = bytes;
  return GALIC_OK;
}

// Arena blocks planned for a launch of n orbits: the running estimate per orbit
// with a margin, and the blocks the lanes may claim and leave unused.
double log_blocks_planned(const galic::Ctx &c, int n) {
  const double lanes = (double)orbit_grid(n, 256, c.fp32_bpc) * 256;
  return (double)n * (c.log_est_blocks + 1.0) * 1.25 + lanes * galic::LOG_CHUNK + 64;
}

constexpr int kLogArenaFull = 1;  // launch_orbit32_log: the arena ran out, plan a smaller launch

constexpr int kLogTableFull = 2;  // orbit32_log_end: an orbit outgrew its block table (log_M doubled)

// One launch of the logging orbit kernel over A.n orbits into the active orbit set's arena, on
// that set's stream, not waited on: the tables sized, the counters cleared, the launch bracketed
// by the set's events.  orbit32_log_end waits for it.
int orbit32_log_begin(galic::Ctx &c, galic::OrbitArgs32 &A) {
  const size_t bb = log_block_bytes();
  const size_t cap = c.slot_bytes / bb;
  if(cap < 64) return set_err(GALIC_ERR_NOMEM, "no room for orbit response logs");
  const size_t nt = (size_t)A.n * c.log_M;
  if(nt > c.log_blocks_len) {
    dfree(c.log_blocks);
    c.log_blocks_len = 0;
    CHK(dalloc(&c.log_blocks, nt));
    c.log_blocks_len = nt;
  }
  if(A.n > c.log_n_len) {
    dfree(c.log_n);
    dfree(c.log_tot);
    c.log_n_len = 0;
    CHK(dalloc(&c.log_n, (size_t)A.n));
    CHK(dalloc(&c.log_tot, (size_t)A.n * galic::kFields));
    c.log_n_len = A.n;
  }
  A.log_blocks = c.log_blocks;
  A.log_n = c.log_n;
  A.log_tot = c.log_tot;
  A.log_M = c.log_M;
  A.log_q = c.log_q = (int *)c.slots;
  A.log_v = c.log_v = (float4 *)((int *)c.slots + cap * galic::LOG_B);
  A.log_next = c.log_next;
  A.log_cap = (unsigned int)std::min<size_t>(cap, 0xf0000000u);  // claims past the cap stay below 2^32
  HIPCHK(hipMemsetAsync(c.log_next, 0, sizeof(unsigned int), c.stream));
  HIPCHK(hipMemsetAsync(c.counter, 0, sizeof(unsigned int), c.stream));
  HIPCHK(hipMemsetAsync(c.orbit_stats, 0, 4 * sizeof(unsigned long long), c.stream));
  HIPCHK(hipEventRecord(c.ev0, c.stream));
  hipLaunchKernelGGL(galic::k_orbit32_log, dim3(orbit_grid(A.n, 256, c.fp32_bpc)), dim3(256), 0, c.stream,
                     fast_model(c), A);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c.ev1, c.stream));
  return GALIC_OK;
}

// Waits for the launch of orbit32_log_begin.  If an orbit needed more blocks than its table
// holds, the table grows and kLogTableFull asks for the same launch again (the orbits and their
// logs do not depend on the launch, so the result is the same); if the arena ran out, the block
// estimate doubles and kLogArenaFull asks the caller for a smaller launch.
int orbit32_log_end(galic::Ctx &c, const galic::OrbitArgs32 &A) {
  const size_t cap = c.slot_bytes / log_block_bytes();
  HIPCHK(hipEventSynchronize(c.ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  unsigned long long st[4];
  unsigned int used = 0;
  HIPCHK(hipMemcpyAsync(st, c.orbit_stats, sizeof(st), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipMemcpyAsync(&used, c.log_next, sizeof(used), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  if(st[3] == 0) {
    c.last_orbit_ms = ms;
    c.acc_orbit_ms += ms;
    c.acc_orbit_launches++;
    c.log_est_blocks = std::max(1.0, (double)std::min<size_t>(used, cap) / A.n);
    return finish_orbit_launch(c, A.n);
  }
  c.acc_log_relaunches++;
  if((size_t)used >= cap) {
    c.log_est_blocks *= 2;
    return kLogArenaFull;
  }
  c.log_M *= 2;  // an orbit needed more blocks than its table holds
  return kLogTableFull;
}

// One launch of the logging orbit kernel, waited on (tables grown and the launch repeated as
// needed); kLogArenaFull if the arena ran out.
int launch_orbit32_log(galic::Ctx &c, galic::OrbitArgs32 A) {
  if(A.n <= 0) return GALIC_OK;
  for(int attempt = 0; attempt < 8; attempt++) {
    CHK(orbit32_log_begin(c, A));
    const int r = orbit32_log_end(c, A);
    if(r != kLogTableFull) return r;
  }
  return set_err(GALIC_ERR_NOMEM, "orbit response logs still overflow after growing the block tables");
}

// Loc += delta in fixed point, and the FP64 copies the global fit reads.
int commit_q(galic::Ctx &c) {
  size_t na = (size_t)GALIC_NTYPES * galic::kFields * c.Ngrid, ni = (size_t)GALIC_NTYPES * 4 * c.Ngrid;
  hipLaunchKernelGGL(galic::k_commit_q, dim3(nblocks(na, 256)), dim3(256), 0, c.stream, c.locA_q, c.locA_dq, c.locA,
                     (size_t)c.Ngrid, galic::kFields, (const double *)c.qinv, na);
  hipLaunchKernelGGL(galic::k_commit_q, dim3(nblocks(ni, 256)), dim3(256), 0, c.stream, c.locI_q, c.locI_dq, c.locI,
                     (size_t)c.Ngrid, 4, (const double *)(c.qinv + 5), ni);
  HIPCHK(hipGetLastError());
  return GALIC_OK;
}

// Bytes the orbit slots of one chunk may take: most of the free HBM (a chunk
// large enough to keep every lane busy until its end), leaving room for
// another context and the library's other buffers.
size_t slot_budget(galic::Ctx &c) {
  size_t fr = 0, tot = 0;
  if(hipMemGetInfo(&fr, &tot) != hipSuccess) fr = (size_t)8 << 30;
  const size_t reserve = (size_t)8 << 30, cap = (size_t)192 << 30;
  size_t have = c.slot_bytes + (fr > reserve ? (fr - reserve) / 4 * 3 : 0);
  return std::max<size_t>((size_t)1 << 30, std::min(cap, have));
}

// The orbits of the initial pass (prepared at o_* [0, n)) through the logging kernel in
// chunks the arena holds, each chunk's logs added to the Loc grids by k_log_to_loc.  The
// first chunk is small: it sets the estimate of log blocks per orbit for the others.
// Adds logs to dst (fixed-point Loc grids [type][5][Ngrid]) through region buckets: the
// entries of each (type, region) are counted on the device, their buckets laid out here, the
// entries scattered, then summed region chunk by region chunk (galic_commit.hip).  A.n items,
// with A.accept: each trial's taken orbit added and its current orbit taken away; without:
// every orbit added.
constexpr int kCommitChunk = 16384;        // bucket entries one workgroup sums
constexpr double kCommitEntries = 1 << 28;  // bucket entries of one initial-pass commit (2 GiB)
int commit_logs(galic::Ctx &c, galic::CommitArgs A, long long *dst) {
  if(A.n <= 0) return GALIC_OK;
  const int gsh = std::min(c.L, galic::kCommitRegionLog2);
  const int nreg = 1 << (c.L - gsh), ng = nreg * nreg, nk = GALIC_NTYPES * ng;
  if(c.cm_nk < nk) {
    dfree(c.cm_cnt);
    dfree(c.cm_start);
    c.cm_nk = 0;
    CHK(dalloc(&c.cm_cnt, 2 * (size_t)nk));
    CHK(dalloc(&c.cm_start, (size_t)nk));
    c.cm_nk = nk;
  }
  A.L = c.L;
  A.gsh = gsh;
  A.log_blocks = c.log_blocks;
  A.log_n = c.log_n;
  A.log_M = c.log_M;
  A.log_q = c.log_q;
  A.log_v = c.log_v;
  for(int f = 0; f < galic::kFields; f++) A.q[f] = c.fp32_scale[f];
  A.gcnt = c.cm_cnt;
  A.gcur = c.cm_cnt + nk;
  HIPCHK(hipMemsetAsync(c.cm_cnt, 0, 2 * sizeof(unsigned int) * nk, c.stream));
  int ncu = 256;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device);
  const int grid = std::min(A.n, ncu * 8);
  const size_t sh = sizeof(unsigned int) * ng;
  hipLaunchKernelGGL(galic::k_commit_count, dim3(grid), dim3(256), sh, c.stream, A);
  HIPCHK(hipGetLastError());
  std::vector<unsigned int> cnt(nk), start(nk);
  HIPCHK(hipMemcpyAsync(cnt.data(), c.cm_cnt, sizeof(unsigned int) * nk, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  size_t total = 0;
  std::vector<int4> chunks;
  for(int k = 0; k < nk; k++) {
    start[k] = (unsigned int)total;
    for(size_t lo = 0; lo < cnt[k]; lo += kCommitChunk)
      chunks.push_back(make_int4(k, (int)(total + lo), (int)(total + std::min<size_t>(cnt[k], lo + kCommitChunk)), 0));
    total += cnt[k];
    if(total > 0x7fffffffu) return set_err(GALIC_ERR_NOMEM, "a commit of %zu region entries exceeds 2^31", total);
  }
  if(total == 0) return GALIC_OK;
  if(c.cm_buf_len < total) {
    dfree(c.cm_buf);
    c.cm_buf_len = 0;
    const size_t want = total + total / 4;
    CHK(dalloc(&c.cm_buf, want));
    c.cm_buf_len = want;
  }
  if(c.cm_chunks_len < chunks.size()) {
    dfree(c.cm_chunks);
    c.cm_chunks_len = 0;
    CHK(dalloc(&c.cm_chunks, chunks.size() * 2));
    c.cm_chunks_len = chunks.size() * 2;
  }
  HIPCHK(hipMemcpyAsync(c.cm_start, start.data(), sizeof(unsigned int) * nk, hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipMemcpyAsync(c.cm_chunks, chunks.data(), sizeof(int4) * chunks.size(), hipMemcpyHostToDevice, c.stream));
  A.gstart = c.cm_start;
  A.gbuf = c.cm_buf;
  hipLaunchKernelGGL(galic::k_commit_scatter, dim3(grid), dim3(256), sh, c.stream, A);
  HIPCHK(hipGetLastError());
  const size_t shr = sizeof(long long) * galic::kFields << (2 * gsh);
  hipLaunchKernelGGL(galic::k_commit_regions, dim3((unsigned int)chunks.size()), dim3(256), shr, c.stream, A,
                     (const int4 *)c.cm_chunks, dst, (const unsigned int *)nullptr);
  HIPCHK(hipGetLastError());
  // (the host vectors of the copies above must outlive them)
  HIPCHK(hipStreamSynchronize(c.stream));
  return GALIC_OK;
}

// Room for a commit of at most cap_entries bucket entries laid out on the device
// (commit_logs_launch); cap_entries <= 2^31 - 1.
int ensure_commit_room(galic::Ctx &c, size_t cap_entries, size_t *cap_chunks) {
  const int gsh = std::min(c.L, galic::kCommitRegionLog2);
  const int nreg = 1 << (c.L - gsh), nk = GALIC_NTYPES * nreg * nreg;
  if(c.cm_nk < nk) {
    dfree(c.cm_cnt);
    dfree(c.cm_start);
    c.cm_nk = 0;
    CHK(dalloc(&c.cm_cnt, 2 * (size_t)nk));
    CHK(dalloc(&c.cm_start, (size_t)nk));
    c.cm_nk = nk;
  }
  if(c.cm_buf_len < cap_entries) {
    dfree(c.cm_buf);
    c.cm_buf_len = 0;
    CHK(dalloc(&c.cm_buf, cap_entries));
    c.cm_buf_len = cap_entries;
  }
  *cap_chunks = (size_t)nk + cap_entries / kCommitChunk + 1;
  if(c.cm_chunks_len < *cap_chunks) {
    dfree(c.cm_chunks);
    c.cm_chunks_len = 0;
    CHK(dalloc(&c.cm_chunks, *cap_chunks));
    c.cm_chunks_len = *cap_chunks;
  }
  if(!c.cm_aux) CHK(dalloc(&c.cm_aux, 2));
  return GALIC_OK;
}

// commit_logs with the bucket layout made on the device (k_commit_layout), nothing waited for
// on the host, so that it can be captured in a graph.  The room (ensure_commit_room) must hold
// the commit's entries; if it does not, cm_aux[1] is set and nothing is committed.
int commit_logs_launch(galic::Ctx &c, galic::CommitArgs A, long long *dst, size_t cap_entries, size_t cap_chunks,
                       int ncu) {
  if(A.n <= 0) return GALIC_OK;
  const int gsh = std::min(c.L, galic::kCommitRegionLog2);
  const int nreg = 1 << (c.L - gsh), ng = nreg * nreg, nk = GALIC_NTYPES * ng;
  A.L = c.L;
  A.gsh = gsh;
  A.log_blocks = c.log_blocks;
  A.log_n = c.log_n;
  A.log_M = c.log_M;
  A.log_q = c.log_q;
  A.log_v = c.log_v;
  for(int f = 0; f < galic::kFields; f++) A.q[f] = c.fp32_scale[f];
  A.gcnt = c.cm_cnt;
  A.gcur = c.cm_cnt + nk;
  A.gstart = c.cm_start;
  A.gbuf = c.cm_buf;
  A.gerr = c.cm_aux + 1;
  hipLaunchKernelGGL(galic::k_fill_u32, dim3(nblocks(2 * (size_t)nk, 256)), dim3(256), 0, c.stream, c.cm_cnt, 0u,
                     2 * (size_t)nk);
  const int grid = std::min(A.n, ncu * 8);
  const size_t sh = sizeof(unsigned int) * ng;
  hipLaunchKernelGGL(galic::k_commit_count, dim3(grid), dim3(256), sh, c.stream, A);
  hipLaunchKernelGGL(galic::k_commit_layout, dim3(1), dim3(1024), 0, c.stream, (const unsigned int *)c.cm_cnt, nk,
                     kCommitChunk, c.cm_start, c.cm_chunks, c.cm_aux, (unsigned int)cap_entries,
                     (unsigned int)cap_chunks, c.cm_aux + 1);
  hipLaunchKernelGGL(galic::k_commit_scatter, dim3(grid), dim3(256), sh, c.stream, A);
  const size_t shr = sizeof(long long) * galic::kFields << (2 * gsh);
  hipLaunchKernelGGL(galic::k_commit_regions, dim3((unsigned int)cap_chunks), dim3(256), shr, c.stream, A,
                     (const int4 *)c.cm_chunks, dst, (const unsigned int *)c.cm_aux);
  HIPCHK(hipGetLastError());
  return GALIC_OK;
}

int fp32_initial_logs(galic::Ctx &c, int n) {
  const size_t bb = log_block_bytes();
  {
    const size_t all = (size_t)(log_blocks_planned(c, n) * bb);
    const size_t one = (size_t)(log_blocks_planned(c, std::min(n, 4096)) * bb);
    CHK(ensure_arena(c, std::min(all, slot_budget(c)), one));
  }
  int a = 0;
  bool first = true;
  while(a < n) {
    const double blocks = (double)(c.slot_bytes / bb);
    const double lanes = (double)orbit_grid(n - a, 256, c.fp32_bpc) * 256;
    double fit = (blocks - 64 - lanes * galic::LOG_CHUNK) / ((c.log_est_blocks + 1.0) * 1.25);
    if(first) fit = std::min(fit, 65536.0);
    const int m = (int)std::max(1.0, std::min((double)(n - a), fit));
    galic::OrbitArgs32 A = orbit_args32(c, m);
    A.ox += a; A.oy += a; A.oz += a;
    A.ovx += a; A.ovy += a; A.ovz += a;
    A.omass += a; A.otint += a; A.otype += a;
    A.tall += a; A.norbit += a; A.status += a; A.steps += a; A.bbox += a;
    const int lr = launch_orbit32_log(c, A);
    if(lr == kLogArenaFull) {
      if(m <= 1) return set_err(GALIC_ERR_NOMEM, "the response log of one orbit does not fit in device memory");
      continue;  // the estimate has doubled: plan the chunk again
    }
    CHK(lr);
    // every record of the chunk's orbits into the Loc grids, in runs of orbits whose bucket
    // entries (about 1.3 per record) stay within kCommitEntries
    const double per_orbit = std::max(1.0, c.log_est_blocks * galic::LOG_B * 1.3);
    const int run = (int)std::max(1.0, std::min((double)m, (double)kCommitEntries / per_orbit));
    for(int b0 = 0; b0 < m; b0 += run) {
      galic::CommitArgs C;
      memset(&C, 0, sizeof(C));
      C.n = std::min(run, m - b0);
      C.orbit_base = b0;
      C.otype = c.o_type + a + b0;
      CHK(commit_logs(c, C, c.locA_q));
    }
    a += m;
    first = false;
  }
  return GALIC_OK;
}

}  // namespace

namespace galic {
// Plain streaming copy: the reference point that calibrates the HBM byte
// counters against a kernel whose traffic is known exactly.
__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for(; i < n; i += stride) dst[i] = src[i];
}

int set_error(int code, const char *msg) { return set_err(code, "%s", msg); }
galic_ctx *bound_ctx() { return g_bound; }
Ctx &ctx_state(galic_ctx *ctx) { return ctx->c; }
}  // namespace galic

// =========================================================================
extern "C" {

const char *galic_last_error(void) { return g_err.c_str(); }
int galic_version(void) { return 10100; }

int galic_copy_probe(int device, size_t nbytes, int reps, double *ms_per_copy) {
  if(nbytes < 16 || reps < 1) return set_err(GALIC_ERR_ARG, "copy probe needs >= 16 bytes and >= 1 rep");
  HIPCHK(hipSetDevice(device));
  size_t n = nbytes / 16;
  float4 *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, n * 16));
  HIPCHK(hipMalloc(&b, n * 16));
  HIPCHK(hipMemset(a, 0, n * 16));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(galic::k_copy, dim3(blocks), dim3(256), 0, 0, (const float4 *)a, b, n);
  HIPCHK(hipEventRecord(e0, 0));
  for(int r = 0; r < reps; r++) hipLaunchKernelGGL(galic::k_copy, dim3(blocks), dim3(256), 0, 0, (const float4 *)a, b, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(e1, 0));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  if(ms_per_copy) *ms_per_copy = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return GALIC_OK;
}

int galic_grid_create(galic_params *p, double outer_frac, double inner_frac) {
  if(!p || !(outer_frac > 0 && outer_frac < 1) || !(inner_frac > 0 && inner_frac < 1))
    return set_err(GALIC_ERR_ARG, "enclosed-mass fractions must lie in (0,1)");
  double fac = outer_frac;
  p->Rmax = p->Halo_A * (fac + sqrt(fac)) / (1 - fac);
  fac = inner_frac;
  double rin = p->Halo_A * (fac + sqrt(fac)) / (1 - fac);
  p->DG_Fac = solve_fac(1 << p->DG_MaxLevel, rin, p->Rmax, &p->DG_Rmin);
  p->FG_Fac = solve_fac(p->FG_Nbin, rin, p->Rmax, &p->FG_Rmin);
  p->EG_Fac = solve_fac(1 << p->EG_MaxLevel, rin, p->Rmax, &p->EG_Rmin);
  if(p->DG_Fac < 0 || p->FG_Fac < 0 || p->EG_Fac < 0)
    return set_err(GALIC_ERR_ARG, "Newton iteration for the grid spacing did not converge");
  return GALIC_OK;
}

int galic_density_geometry(const galic_params *p, double *cellsize, double *cellvol, double *distance) {
  if(!p || p->DG_MaxLevel < 1 || p->DG_MaxLevel > 12) return set_err(GALIC_ERR_ARG, "bad DG_MaxLevel");
  const int L = p->DG_MaxLevel, nb = 1 << L, Ns = GALIC_NSTACK(L);
  if(distance) {
    std::vector<double> logR(Ns, 0.0), logZ(Ns, 0.0);
    for(int k = 0; k < nb; k++)
      for(int j = 0; j < nb; j++) {
        double z = p->DG_Rmin * 0.5 * (pow(p->DG_Fac, k) + pow(p->DG_Fac, k + 1) - 2.0);
        double R = p->DG_Rmin * 0.5 * (pow(p->DG_Fac, j) + pow(p->DG_Fac, j + 1) - 2.0);
        logR[GALIC_STACKOFFSET(L, k, j)] = log(R);
        logZ[GALIC_STACKOFFSET(L, k, j)] = log(z);
      }
    smooth_stack_host(logR.data(), L);
    smooth_stack_host(logZ.data(), L);
    for(int l = L - 1; l >= 0; l--) {
      int n = 1 << l, f = 1 << (L - l);
      for(int i = 0; i < n; i++)
        for(int j = 0; j < n; j++) {
          logR[GALIC_STACKOFFSET(l, i, j)] /= (f * f);
          logZ[GALIC_STACKOFFSET(l, i, j)] /= (f * f);
        }
    }
    for(int l = L; l >= 0; l--) {
      int n = 1 << l;
      for(int i = 0; i < n; i++)
        for(int j = 0; j < n; j++)
          distance[GALIC_STACKOFFSET(l, i, j)] =
              sqrt(pow(exp(logR[GALIC_STACKOFFSET(l, i, j)]), 2) + pow(exp(logZ[GALIC_STACKOFFSET(l, i, j)]), 2));
    }
  }
  for(int k = 0; k < nb; k++)
    for(int j = 0; j < nb; j++) {
      int i = k * nb + j;
      double r1 = p->DG_Rmin * (pow(p->DG_Fac, j) - 1.0);
      double r2 = p->DG_Rmin * (pow(p->DG_Fac, j + 1) - 1.0);
      double z1 = p->DG_Rmin * (pow(p->DG_Fac, k) - 1.0);
      double z2 = p->DG_Rmin * (pow(p->DG_Fac, k + 1) - 1.0);
      double vol = M_PI * (r2 * r2 - r1 * r1) * (z2 - z1);
      vol *= 2;
      if(cellsize) cellsize[i] = (r2 - r1 > z2 - z1) ? r2 - r1 : z2 - z1;
      if(cellvol) cellvol[i] = vol;
    }
  return GALIC_OK;
}

int galic_create(const galic_params *p, int device, galic_ctx **out) {
  if(!p || !out) return set_err(GALIC_ERR_ARG, "null argument");
  *out = nullptr;
  if(p->DG_MaxLevel != p->EG_MaxLevel)
    return set_err(GALIC_ERR_ARG, "DG_MaxLevel (%d) must equal EG_MaxLevel (%d)", p->DG_MaxLevel, p->EG_MaxLevel);
  if(p->DG_MaxLevel < 2 || p->DG_MaxLevel > 9)
    return set_err(GALIC_ERR_ARG, "DG_MaxLevel %d outside the supported range 2..9", p->DG_MaxLevel);
  if(p->FG_Nbin < 2) return set_err(GALIC_ERR_ARG, "FG_Nbin must be at least 2");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if(device < 0 || device >= ndev) return set_err(GALIC_ERR_ARG, "device %d not present (%d devices)", device, ndev);
  HIPCHK(hipSetDevice(device));

  galic_ctx *ctx = new galic_ctx();
  galic::Ctx &c = ctx->c;
  c.p = *p;
  c.device = device;
  c.L = p->DG_MaxLevel;
  c.Nbin = 1 << c.L;
  c.Ngrid = c.Nbin * c.Nbin;
  c.Nstack = GALIC_NSTACK(c.L);
  c.FG_Ngrid = p->FG_Nbin * p->FG_Nbin;
  HIPCHK(hipStreamCreate(&c.stream));
  HIPCHK(hipEventCreate(&c.ev0));
  HIPCHK(hipEventCreate(&c.ev1));

  CHK(dalloc(&c.FG_Pot, c.FG_Ngrid));
  CHK(dalloc(&c.FG_DPotDR, c.FG_Ngrid));
  CHK(dalloc(&c.FG_DPotDz, c.FG_Ngrid));
  CHK(dalloc(&c.FG_Vstream, (size_t)GALIC_NTYPES * c.FG_Ngrid));
  CHK(dalloc(&c.FG_Acc_f, c.FG_Ngrid));
  CHK(dalloc(&c.FG_Vstream_f, (size_t)GALIC_NTYPES * c.FG_Ngrid));
  c.vstream_nz = 0;
  CHK(dalloc(&c.DG_CellSize, c.Ngrid));
  CHK(dalloc(&c.DG_CellSize_f, c.Ngrid));
  CHK(dalloc(&c.DG_Cell1_f, (size_t)c.Nbin));
  c.cell1_ok = false;
  CHK(dalloc(&c.stacks, (size_t)GALIC_NTYPES * galic::ST_COUNT * c.Nstack));
  CHK(dalloc(&c.weights, (size_t)GALIC_NTYPES * galic::W_COUNT * c.Nstack));
  CHK(dalloc(&c.merit_bw, (size_t)GALIC_NTYPES * galic::kFields * c.Nstack));
  CHK(dalloc(&c.merit_bwf, (size_t)GALIC_NTYPES * galic::kFields * c.Nstack));
  CHK(dalloc(&c.cutB_of, (size_t)GALIC_NTYPES * c.Ngrid));
  CHK(dalloc(&c.locA, (size_t)GALIC_NTYPES * galic::kFields * c.Ngrid));
  CHK(dalloc(&c.locA_d, (size_t)GALIC_NTYPES * galic::kFields * c.Ngrid));
  CHK(dalloc(&c.locI, (size_t)GALIC_NTYPES * 4 * c.Ngrid));
  CHK(dalloc(&c.locI_d, (size_t)GALIC_NTYPES * 4 * c.Ngrid));
  CHK(dalloc(&c.locM, (size_t)GALIC_NTYPES * c.Ngrid));
  CHK(dalloc(&c.locA_q, (size_t)GALIC_NTYPES * galic::kFields * c.Ngrid));
  CHK(dalloc(&c.locA_dq, (size_t)GALIC_NTYPES * galic::kFields * c.Ngrid));
  CHK(dalloc(&c.locI_q, (size_t)GALIC_NTYPES * 4 * c.Ngrid));
  CHK(dalloc(&c.locI_dq, (size_t)GALIC_NTYPES * 4 * c.Ngrid));
  CHK(dalloc(&c.locM_q, (size_t)GALIC_NTYPES * c.Ngrid));
  CHK(dalloc(&c.qinv, 9));
  CHK(dalloc(&c.tovs_dev, GALIC_NTYPES));
  CHK(dalloc(&c.fit, (size_t)GALIC_NTYPES * 16));
  CHK(dalloc(&c.fit_part, (size_t)3 * galic::kFitBlocks * 5));
  CHK(dalloc(&c.counter, 1));
  CHK(dalloc(&c.orbit_stats, 4));
  CHK(dalloc(&c.log_next, 1));

  // geometry
  ctx->distance.assign(c.Nstack, 0.0);
  ctx->cellsize.assign(c.Ngrid, 0.0);
  CHK(galic_density_geometry(p, ctx->cellsize.data(), nullptr, ctx->distance.data()));
  HIPCHK(hipMemcpy(c.DG_CellSize, ctx->cellsize.data(), sizeof(double) * c.Ngrid, hipMemcpyHostToDevice));
  CHK(upload_cell_sizes_f(c, ctx->cellsize.data()));
  for(int t = 1; t < GALIC_NTYPES; t++)
    for(int k = 0; k < 6; k++) ctx->targets[t][k].assign(c.Nstack, 0.0);

  // device model
  galic::DevModel &M = c.M;
  memset(&M, 0, sizeof(M));
  M.G = p->G; M.V200 = p->V200; M.R200 = p->R200; M.M200 = p->M200;
  M.Halo_Mass = p->Halo_Mass; M.Halo_A = p->Halo_A;
  M.Bulge_Mass = p->Bulge_Mass; M.Bulge_A = p->Bulge_A;
  M.halo_offgrid_fac = (p->Halo_Mass > 0) ? p->M200 / p->Halo_Mass : 0.0;
  M.use_table = (p->Disk_Mass > 0 || (p->Halo_Mass > 0 && p->SampleForceNhalo > 0) ||
                 (p->Bulge_Mass > 0 && p->SampleForceNbulge > 0));
  M.add_halo = (p->Halo_Mass > 0 && p->SampleForceNhalo == 0);
  M.add_bulge = (p->Bulge_Mass > 0 && p->SampleForceNbulge == 0);
  for(int t = 0; t < GALIC_NTYPES; t++) M.tovs[t] = p->TypeOfVelocityStructure[t];
  M.fo = p->TimeStepFactorOrbit;
  M.fc = p->TimeStepFactorCellCross;
  M.max_steps = p->MaxStepsPerOrbit > 0 ? p->MaxStepsPerOrbit : 100000000LL;
  M.FG_Nbin = p->FG_Nbin; M.FG_Rmin = p->FG_Rmin; M.FG_logFac = log(p->FG_Fac);
  M.DG_Nbin = c.Nbin; M.DG_L = c.L; M.DG_Rmin = p->DG_Rmin; M.DG_logFac = log(p->DG_Fac);
  M.EG_Nbin = 1 << p->EG_MaxLevel; M.EG_L = p->EG_MaxLevel; M.EG_Rmin = p->EG_Rmin; M.EG_logFac = log(p->EG_Fac);
  M.fFG_invRmin = (float)(1.0 / p->FG_Rmin); M.fFG_invlogFac = (float)(1.0 / log(p->FG_Fac));
  M.fDG_invRmin = (float)(1.0 / p->DG_Rmin); M.fDG_invlogFac = (float)(1.0 / log(p->DG_Fac));
  M.FG_Pot = c.FG_Pot; M.FG_DPotDR = c.FG_DPotDR; M.FG_DPotDz = c.FG_DPotDz;
  for(int t = 0; t < GALIC_NTYPES; t++) {
    M.FG_Vstream[t] = c.FG_Vstream + (size_t)t * c.FG_Ngrid;
    M.FG_Vstream_f[t] = c.FG_Vstream_f + (size_t)t * c.FG_Ngrid;
  }
  M.DG_CellSize = c.DG_CellSize;
  M.DG_CellSize_f = c.DG_CellSize_f;
  M.FG_Acc_f = c.FG_Acc_f;
  // fixed-point scales of the FP32 path: 2^-40 in units of mass, of mass*V200^2 for moments
  c.fp32_scale[0] = ldexp(1.0, 40);
  for(int f = 1; f < 5; f++) c.fp32_scale[f] = ldexp(1.0, 40) / (p->V200 * p->V200);
  c.fp32_iscale = ldexp(1.0, 40) / (p->V200 * p->V200);
  {
    double inv[9];
    for(int f = 0; f < 5; f++) inv[f] = 1.0 / c.fp32_scale[f];
    for(int f = 0; f < 4; f++) inv[5 + f] = 1.0 / c.fp32_iscale;
    HIPCHK(hipMemcpy(c.qinv, inv, sizeof(inv), hipMemcpyHostToDevice));
    int tv[GALIC_NTYPES];
    for(int t = 0; t < GALIC_NTYPES; t++) tv[t] = p->TypeOfVelocityStructure[t];
    HIPCHK(hipMemcpy(c.tovs_dev, tv, sizeof(tv), hipMemcpyHostToDevice));
  }
  *out = ctx;
  return GALIC_OK;
}

int galic_destroy(galic_ctx *ctx) {
  if(!ctx) return GALIC_OK;
  galic::Ctx &c = ctx->c;
  (void)hipSetDevice(c.device);
  (void)hipStreamSynchronize(c.stream);
  dfree(c.FG_Pot); dfree(c.FG_DPotDR); dfree(c.FG_DPotDz); dfree(c.FG_Vstream);
  dfree(c.FG_Acc_f); dfree(c.FG_Vstream_f); dfree(c.DG_CellSize); dfree(c.DG_CellSize_f); dfree(c.DG_Cell1_f);
  dfree(c.stacks); dfree(c.weights); dfree(c.cutB_of); dfree(c.merit_bw); dfree(c.merit_bwf); dfree(c.tcut);
  dfree(c.locA); dfree(c.locA_d); dfree(c.locI); dfree(c.locI_d); dfree(c.locM);
  dfree(c.locA_q); dfree(c.locA_dq); dfree(c.locI_q); dfree(c.locI_dq);
  dfree(c.fit); dfree(c.locM_q); dfree(c.qinv); dfree(c.tovs_dev); dfree(c.t_vk); dfree(c.t_dFk); dfree(c.t_vinst);
  galic::Particles &P = c.P;
  dfree(P.x); dfree(P.y); dfree(P.z); dfree(P.vx); dfree(P.vy); dfree(P.vz);
  dfree(P.mass); dfree(P.tint); dfree(P.vesc); dfree(P.vr2); dfree(P.vt2); dfree(P.vp2); dfree(P.vq2);
  dfree(P.type); dfree(P.orbits); dfree(P.disp);
  dfree(c.t_idx); dfree(c.t_v); dfree(c.t_accept); dfree(c.t_dF); dfree(c.t_best); dfree(c.psteps);
  for(int t = 0; t < GALIC_NTYPES; t++) dfree(c.disp_tab[t]);
  dfree(c.cs_scratch);
  dfree(c.g32_scratch);
  dfree(c.tile_list);
  dfree(c.tile_aux);
  dfree(c.cm_cnt); dfree(c.cm_start); dfree(c.cm_chunks); dfree(c.cm_buf); dfree(c.cm_aux); dfree(c.cm_bound);
  dfree(c.merit_prof);
  dfree(c.lookup_buf);
  dfree(c.fit_part);
  dfree(c.tsrc);
  if(c.alt_ready) {
    swap_orbit_set(c);
    free_orbit_set(c);
    swap_orbit_set(c);
  }
  free_orbit_set(c);
  if(g_bound == ctx) g_bound = nullptr;
  delete ctx;
  return GALIC_OK;
}

int galic_set_force_tables(galic_ctx *ctx, const double *pot, const double *dpotdr, const double *dpotdz) {
  if(!ctx || !pot || !dpotdr || !dpotdz) return set_err(GALIC_ERR_ARG, "null argument");
  galic::Ctx &c = ctx->c;
  HIPCHK(hipSetDevice(c.device));
  size_t n = c.FG_Ngrid;
  HIPCHK(hipMemcpy(c.FG_Pot, pot, n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.FG_DPotDR, dpotdr, n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.FG_DPotDz, dpotdz, n * sizeof(double), hipMemcpyHostToDevice));
  std::vector<float2> f(n);
  for(size_t i = 0; i < n; i++) f[i] = make_float2((float)dpotdr[i], (float)dpotdz[i]);
  HIPCHK(hipMemcpy(c.FG_Acc_f, f.data(), n * sizeof(float2), hipMemcpyHostToDevice));
  return GALIC_OK;
}

int galic_get_force_tables(galic_ctx *ctx, double *pot, double *dpotdr, double *dpotdz) {
  if(!ctx) return set_err(GALIC_ERR_ARG, "null context");
  galic::Ctx &c = ctx->c;
  HIPCHK(hipSetDevice(c.device));
  const size_t n = c.FG_Ngrid;
  if(pot) HIPCHK(hipMemcpy(pot, c.FG_Pot, n * sizeof(double), hipMemcpyDeviceToHost));
  if(dpotdr) HIPCHK(hipMemcpy(dpotdr, c.FG_DPotDR, n * sizeof(double), hipMemcpyDeviceToHost));
  if(dpotdz) HIPCHK(hipMemcpy(dpotdz, c.FG_DPotDz, n * sizeof(double), hipMemcpyDeviceToHost));
  return GALIC_OK;
}

int galic_sample_tree_sources(galic_ctx *ctx, unsigned int seed) {
  if(!ctx) return set_err(GALIC_ERR_ARG, "null context");
  galic::Ctx &c = ctx->c;
  HIPCHK(hipSetDevice(c.device));
  // grid.c:659-678: SampleForceN{halo,disk,bulge} fresh points of each sampled component,
  // each of mass M / N (halo.c:17-40, disk.c:17-51, bulge.c:17-38)
  const int counts[GALIC_NTYPES] = {0, c.p.SampleForceNhalo, c.p.SampleForceNdisk, c.p.SampleForceNbulge};
  long long total = 0;
  for(int t = 1; t < GALIC_NTYPES; t++)
    if(c.p.MType[t] > 0 && counts[t] > 0) total += counts[t];
  if(total < 1) return set_err(GALIC_ERR_STATE, "no component is sampled (SampleForceN of every component is 0)");
  if(total > 0x7fffffff) return set_err(GALIC_ERR_ARG, "%lld sources", total);
  std::vector<double> pos(3 * (size_t)total), mass(total);
  double *dpts = nullptr;
  unsigned long long *cnt = nullptr;
  size_t at = 0;
  int rc = GALIC_OK;
  for(int t = 1; t < GALIC_NTYPES && rc == GALIC_OK; t++) {
    if(!(c.p.MType[t] > 0) || counts[t] <= 0) continue;
    const int n = counts[t];
    CHK(dalloc(&dpts, 3 * (size_t)n));
    CHK(dalloc(&cnt, 2));
    galic::TargetArgs A;
    memset(&A, 0, sizeof(A));
    A.M = c.M;
    A.type = t;
    A.tovs = c.p.TypeOfVelocityStructure[t];
    A.n = n;
    A.seed = seed;
    A.stream = 200 + t;
    A.Rmax = c.p.Rmax;
    A.Disk_H = c.p.Disk_H;
    A.Disk_Z0 = c.p.Disk_Z0;
    A.stretch = 1.0;
    A.pos_out = dpts;
    A.counts = cnt;
    A.count = 1;
    if(galic::launch_sample_targets(A, orbit_grid(n, 256, 8), 0, c.stream) != 0)
      rc = set_err(GALIC_ERR_HIP, "source sampling launch failed");
    unsigned long long hc[2] = {0, 0};
    if(rc == GALIC_OK && (hipMemcpy(pos.data() + 3 * at, dpts, sizeof(double) * 3 * n, hipMemcpyDeviceToHost) != hipSuccess ||
                          hipMemcpy(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost) != hipSuccess))
      rc = set_err(GALIC_ERR_HIP, "copying the sampled sources failed");
    if(rc == GALIC_OK && hc[1] != 0) rc = set_err(GALIC_ERR_ARG, "%llu sources of type %d found no point", hc[1], t);
    for(int i = 0; i < n; i++) mass[at + i] = c.p.MType[t] / n;
    at += n;
    dfree(dpts);
    dfree(cnt);
  }
  if(rc != GALIC_OK) return rc;
  return galic_set_tree_sources(ctx, (int)total, pos.data(), mass.data());
}

void forcedensitygrid_create(void) {
  // grid.c:329-487 for the bound context: the density-grid geometry (DG_CellSize, DGs_Distance)
  // from its parameters, which hold the Fac and Rmin of the grids (galic_grid_create)
  galic_ctx *ctx = galic::bound_ctx();
  if(!ctx) {
    fprintf(stderr, "galic drop-in forcedensitygrid_create failed: no context bound\n");
    exit(1);
  }
  galic::Ctx &c = ctx->c;
  std::vector<double> cs(c.Ngrid), dist(c.Nstack);
  bool ok = galic_density_geometry(&c.p, cs.data(), nullptr, dist.data()) == GALIC_OK &&
            hipSetDevice(c.device) == hipSuccess &&
            hipMemcpy(c.DG_CellSize, cs.data(), sizeof(double) * c.Ngrid, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && upload_cell_sizes_f(c, cs.data()) == GALIC_OK;
  for(int t = 0; t < GALIC_NTYPES && ok; t++)
    ok = hipMemcpy(c.stacks + ((size_t)t * galic::ST_COUNT + galic::ST_DIST) * c.Nstack, dist.data(),
                   sizeof(double) * c.Nstack, hipMemcpyHostToDevice) == hipSuccess;
  if(!ok) {
    fprintf(stderr, "galic drop-in forcedensitygrid_create failed: %s\n", galic_last_error());
    exit(1);
  }
  ctx->cellsize = cs;
  ctx->distance = dist;
}

void forcedensitygrid_calculate(void) {
  // grid.c:621-757 for the bound context: its tree sources (drawn on the device from the
  // sampled components, seed 42 as init.c:25-27 seeds rank 0, if none were set), the
  // reference's walk (FP64, monopole, spline softening), the tables installed in place
  galic_ctx *ctx = galic::bound_ctx();
  int rc = ctx ? GALIC_OK : set_err(GALIC_ERR_STATE, "no context bound");
  if(rc == GALIC_OK && !ctx->c.tsrc) rc = galic_sample_tree_sources(ctx, 42);
  if(rc == GALIC_OK) rc = galic_retabulate(ctx, 0, nullptr);
  if(rc != GALIC_OK) {
    fprintf(stderr, "galic drop-in forcedensitygrid_calculate failed: %s\n", galic_last_error());
    exit(1);
  }
}

int galic_set_vstream(galic_ctx *ctx, int type, const double *vstream) {
  if(!ctx || type < 1 || type > 3) return set_err(GALIC_ERR_ARG, "bad type %d", type);
  galic::Ctx &c = ctx->c;
  HIPCHK(hipSetDevice(c.device));
  size_t n = c.FG_Ngrid;
  std::vector<double> v(n, 0.0);
  if(vstream) memcpy(v.data(), vstream, n * sizeof(double));
  HIPCHK(hipMemcpy(c.FG_Vstream + (size_t)type * n, v.data(), n * sizeof(double), hipMemcpyHostToDevice));
  std::vector<float> f(n);
  bool nz = false;
  for(size_t i = 0; i < n; i++) {
    f[i] = (float)v[i];
    nz = nz || f[i] != 0.f;
  }
  HIPCHK(hipMemcpy(c.FG_Vstream_f + (size_t)type * n, f.data(), n * sizeof(float), hipMemcpyHostToDevice));
  c.vstream_nz = nz ? (c.vstream_nz | (1u << type)) : (c.vstream_nz & ~(1u << type));
  return GALIC_OK;
}

int galic_set_targets(galic_ctx *ctx, int type, const double *dgs_mass_target, const double *egs_mass_target,
                      const double *egs_r, const double *egs_t, const double *egs_p, const double *egs_q) {
  if(!ctx || type < 1 || type > 3) return set_err(GALIC_ERR_ARG, "bad type %d", type);
  if(!dgs_mass_target || !egs_mass_target || !egs_r || !egs_t || !egs_p || !egs_q)
    return set_err(GALIC_ERR_ARG, "null target stack");
  galic::Ctx &c = ctx->c;
  HIPCHK(hipSetDevice(c.device));
  const double *src[6] = {dgs_mass_target, egs_mass_target, egs_r, egs_t, egs_p, egs_q};
  const int dst[6] = {galic::ST_MT, galic::ST_EMT, galic::ST_ER, galic::ST_ET, galic::ST_EP, galic::ST_EQ};
  for(int k = 0; k < 6; k++) {
    ctx->targets[type][k].assign(src[k], src[k] + c.Nstack);
    HIPCHK(hipMemcpy(c.stacks + ((size_t)type * galic::ST_COUNT + dst[k]) * c.Nstack, src[k],
                     sizeof(double) * c.Nstack, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(c.stacks + ((size_t)type * galic::ST_COUNT + galic::ST_DIST) * c.Nstack, ctx->distance.data(),
                   sizeof(double) * c.Nstack, hipMemcpyHostToDevice));
  c.targets_set[type] = true;
  c.cuts_ready = false;
  return GALIC_OK;
}

int galic_set_disp_tables(galic_ctx *ctx, int type, double beta, const double *disp_r, const double *disp_z,
                          const double *disp_phi, const double *tilt_r, const double *tilt_theta) {
  if(!ctx || type < 1 || type > 3) return set_err(GALIC_ERR_ARG, "bad type %d", type);
  galic::Ctx &c = ctx->c;
  const int tovs = c.p.TypeOfVelocityStructure[type];
  const bool need[5] = {tovs <= 1, tovs == 2, tovs >= 2, tovs == 3, tovs == 3};
  const double *src[5] = {disp_r, disp_z, disp_phi, tilt_r, tilt_theta};
  const char *names[5] = {"disp_r", "disp_z", "disp_phi", "tilt_r", "tilt_theta"};
  for(int k = 0; k < 5; k++)
    if(need[k] && !src[k])
      return set_err(GALIC_ERR_ARG, "velocity structure %d of type %d needs the %s table", tovs, type, names[k]);
  HIPCHK(hipSetDevice(c.device));
  const size_t nb = c.p.FG_Nbin, n2 = nb * nb, total = nb + 4 * n2;
  std::vector<double> h(total, 0.0);
  double bound = 0;
  size_t off = 0;
  for(int k = 0; k < 5; k++) {
    const size_t len = k == 0 ? nb : n2;
    if(src[k])
      for(size_t i = 0; i < len; i++) {
        if(!std::isfinite(src[k][i])) return set_err(GALIC_ERR_ARG, "%s[%zu] is not finite", names[k], i);
        h[off + i] = src[k][i];
        bound = std::max(bound, fabs(src[k][i]));
      }
    off += len;
  }
  if(!c.disp_tab[type]) CHK(dalloc(&c.disp_tab[type], total));
  HIPCHK(hipMemcpy(c.disp_tab[type], h.data(), total * sizeof(double), hipMemcpyHostToDevice));
  c.disp_beta[type] = beta;
  // (1 - beta) scales the tangential dispersions of structure 1; a slope-dependent beta lies in [0.05, 0.65]
  c.disp_bound[type] = bound * (1 + (beta < 1 ? fabs(beta) : 1.0));
  return GALIC_OK;
}

int galic_sample_targets(galic_ctx *ctx, int type, long long nsamples, unsigned int seed, unsigned int stream,
                         double stretch, double *dgs_mass_target, double *egs_targets, double *points,
                         long long *nbinned, float *kernel_ms) {
  if(!ctx || type < 1 || type > 3) return set_err(GALIC_ERR_ARG, "bad type %d", type);
  galic::Ctx &c = ctx->c;
  if(!dgs_mass_target && !egs_targets) return set_err(GALIC_ERR_ARG, "no target requested");
  // int64 sums of up to 2^32 units per point stay below 2^62
  if(nsamples < 1 || nsamples > (1ll << 30)) return set_err(GALIC_ERR_ARG, "nsamples must be in [1, 2^30]");
  if(!(stretch > 0)) return set_err(GALIC_ERR_ARG, "stretch must be positive");
  if(!(c.p.MType[type] > 0)) return set_err(GALIC_ERR_ARG, "type %d has no mass", type);
  if(egs_targets && !c.disp_tab[type])
    return set_err(GALIC_ERR_ARG, "the dispersion targets of type %d need galic_set_disp_tables first", type);
  HIPCHK(hipSetDevice(c.device));
  const int tovs = c.p.TypeOfVelocityStructure[type];
  const size_t ng = c.Ngrid, nb = c.p.FG_Nbin, n2 = nb * nb;
  long long *maps = nullptr;
  unsigned long long *counts = nullptr;
  CHK(dalloc(&maps, 6 * ng));
  CHK(dalloc(&counts, 2));
  galic::TargetArgs A;
  memset(&A, 0, sizeof(A));
  A.M = c.M;
  A.type = type;
  A.tovs = tovs;
  A.n = nsamples;
  A.seed = seed;
  A.stream = stream;
  A.Rmax = c.p.Rmax;
  A.Disk_H = c.p.Disk_H;
  A.Disk_Z0 = c.p.Disk_Z0;
  A.stretch = stretch;
  A.beta = c.disp_beta[type];
  if(c.disp_tab[type]) {
    A.disp_r = c.disp_tab[type];
    A.disp_z = A.disp_r + nb;
    A.disp_phi = A.disp_z + n2;
    A.tilt_r = A.disp_phi + n2;
    A.tilt_theta = A.tilt_r + n2;
  }
  A.dg = dgs_mass_target ? maps : nullptr;
  A.eg = egs_targets ? maps + ng : nullptr;
  A.qdg = ldexp(1.0, 32);
  A.qeg[0] = ldexp(1.0, 32);
  if(egs_targets) {
    // every dispersion value stays below 2^e: the q field adds the squared streaming velocity
    std::vector<double> v(n2);
    HIPCHK(hipMemcpy(v.data(), c.FG_Vstream + (size_t)type * n2, n2 * sizeof(double), hipMemcpyDeviceToHost));
    double vmax = 0;
    for(double x : v) vmax = std::max(vmax, fabs(x));
    int e = 0;
    frexp(std::max(c.disp_bound[type] + vmax * vmax, 1e-300), &e);
    for(int f = 1; f < 5; f++) A.qeg[f] = ldexp(1.0, 32 - e);
  }
  A.counts = counts;
  double *dpts = nullptr;
  if(points) {
    CHK(dalloc(&dpts, 3 * (size_t)nsamples));
    A.pos_out = dpts;
  }
  const int grid = orbit_grid((int)std::min<long long>(nsamples, 1 << 30), 256, 8);
  // The maps are summed per workgroup in LDS where they fit in target_lds bytes (64^2
  // grids: the density map, then the five energy maps two at a time), so the
  // atomics on the small maps in memory are few; each launch draws the same points
  // again (cheap beside the atomics), and only the first counts them and writes them out.
  // (64 KiB keeps two workgroups per CU; less where the device gives less)
  size_t target_lds = 65536;
  {
    int v = 0;
    if(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, c.device) == hipSuccess && v > 0)
      target_lds = std::min(target_lds, (size_t)v);
  }
  const size_t per_map = ng * sizeof(long long), fit = target_lds / per_map;
  struct Pass {
    bool dg;
    int f0, f1;
  };
  std::vector<Pass> passes;
  if(fit == 0) {
    passes.push_back({dgs_mass_target != nullptr, 0, egs_targets ? 5 : 0});
  } else {
    if(dgs_mass_target) passes.push_back({true, 0, 0});
    if(egs_targets)
      for(int f0 = 0; f0 < 5; f0 += (int)fit) passes.push_back({false, f0, std::min(5, f0 + (int)fit)});
  }
  HIPCHK(hipEventRecord(c.ev0, c.stream));
  for(size_t i = 0; i < passes.size(); i++) {
    galic::TargetArgs P = A;
    P.dg = passes[i].dg ? A.dg : nullptr;
    P.f0 = passes[i].f0;
    P.f1 = passes[i].f1;
    P.lds = fit > 0 ? 1 : 0;
    P.count = i == 0 ? 1 : 0;
    P.pos_out = i == 0 ? A.pos_out : nullptr;
    const size_t shmem = P.lds ? per_map * ((P.dg ? 1 : 0) + (size_t)(P.f1 - P.f0)) : 0;
    if(galic::launch_sample_targets(P, grid, shmem, c.stream) != 0) {
      dfree(maps);
      dfree(counts);
      dfree(dpts);
      return set_err(GALIC_ERR_HIP, "target sampling launch failed");
    }
  }
  HIPCHK(hipEventRecord(c.ev1, c.stream));
  HIPCHK(hipEventSynchronize(c.ev1));
  HIPCHK(hipEventElapsedTime(&c.last_target_ms, c.ev0, c.ev1));
  std::vector<long long> hq(6 * ng);
  unsigned long long hc[2];
  HIPCHK(hipMemcpy(hq.data(), maps, 6 * ng * sizeof(long long), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(hc, counts, sizeof(hc), hipMemcpyDeviceToHost));
  if(points) HIPCHK(hipMemcpy(points, dpts, 3 * (size_t)nsamples * sizeof(double), hipMemcpyDeviceToHost));
  dfree(maps);
  dfree(counts);
  dfree(dpts);
  if(hc[1] != 0) return set_err(GALIC_ERR_ARG, "%llu samples found no point inside Rmax", hc[1]);
  if(hc[0] == 0) return set_err(GALIC_ERR_ARG, "no sample fell inside the grid");
  // normalisation to the component's mass (grid.c:603-611, 1051-1062), then the pyramid
  const double norm = c.p.MType[type] / (double)hc[0];
  const size_t off = GALIC_STACKOFFSET(c.L, 0, 0);
  std::vector<double> st(c.Nstack);
  auto finish = [&](const long long *q, double unit, double *out) {
    std::fill(st.begin(), st.end(), 0.0);
    for(size_t i = 0; i < ng; i++) st[off + i] = (double)q[i] / unit * norm;
    smooth_stack_host(st.data(), c.L);
    memcpy(out, st.data(), sizeof(double) * c.Nstack);
  };
  if(dgs_mass_target) finish(hq.data(), A.qdg, dgs_mass_target);
  if(egs_targets)
    for(int f = 0; f < 5; f++) finish(hq.data() + (1 + f) * ng, A.qeg[f], egs_targets + (size_t)f * c.Nstack);
  if(nbinned) *nbinned = (long long)hc[0];
  if(kernel_ms) *kernel_ms = c.last_target_ms;
  return GALIC_OK;
}

int galic_set_particles(galic_ctx *ctx, int n, const double *pos, const double *vel, const double *mass,
                        const double *tint, const int *type, const double *vesc) {
  if(!ctx || n <= 0 || !pos || !vel || !mass || !tint || !type) return set_err(GALIC_ERR_ARG, "bad particle input");
  galic::Ctx &c = ctx->c;
  HIPCHK(hipSetDevice(c.device));
  for(int i = 0; i < n; i++) {
    if(type[i] < 1 || type[i] > 3) return set_err(GALIC_ERR_ARG, "particle %d: unknown type %d", i, type[i]);
    if(!type_present(c.p, type[i]))
      return set_err(GALIC_ERR_ARG, "particle %d has type %d but NType[%d] is 0", i, type[i], type[i]);
  }
  galic::Particles &P = c.P;
  dfree(P.x); dfree(P.y); dfree(P.z); dfree(P.vx); dfree(P.vy); dfree(P.vz);
  dfree(P.mass); dfree(P.tint); dfree(P.vesc); dfree(P.vr2); dfree(P.vt2); dfree(P.vp2); dfree(P.vq2);
  dfree(P.type); dfree(P.orbits); dfree(P.disp); dfree(c.psteps);
  P.n = n;
  CHK(dalloc(&c.psteps, n));
  CHK(dalloc(&P.x, n)); CHK(dalloc(&P.y, n)); CHK(dalloc(&P.z, n));
  CHK(dalloc(&P.vx, n)); CHK(dalloc(&P.vy, n)); CHK(dalloc(&P.vz, n));
  CHK(dalloc(&P.mass, n)); CHK(dalloc(&P.tint, n)); CHK(dalloc(&P.vesc, n));
  CHK(dalloc(&P.vr2, n)); CHK(dalloc(&P.vt2, n)); CHK(dalloc(&P.vp2, n)); CHK(dalloc(&P.vq2, n));
  CHK(dalloc(&P.type, n)); CHK(dalloc(&P.orbits, n));
  c.max_mass = 0;
  c.max_vesc = 0;
  for(int i = 0; i < n; i++) {
    c.max_mass = std::max(c.max_mass, mass[i]);
    c.max_vesc = std::max(c.max_vesc, vesc[i]);
  }
  std::vector<double> tmp(n);
  double *dst3[3] = {P.x, P.y, P.z};
  for(int k = 0; k < 3; k++) {
    for(int i = 0; i < n; i++) tmp[i] = pos[3 * i + k];
    HIPCHK(hipMemcpy(dst3[k], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  }
  double *dstv[3] = {P.vx, P.vy, P.vz};
  for(int k = 0; k < 3; k++) {
    for(int i = 0; i < n; i++) tmp[i] = vel[3 * i + k];
    HIPCHK(hipMemcpy(dstv[k], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(P.mass, mass, sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P.tint, tint, sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P.type, type, sizeof(int) * n, hipMemcpyHostToDevice));
  if(vesc) HIPCHK(hipMemcpy(P.vesc, vesc, sizeof(double) * n, hipMemcpyHostToDevice));
  ctx->initial_done = false;
  ctx->fp32_initial_done = false;
  return GALIC_OK;
}

int galic_get_particles(galic_ctx *ctx, double *vel, int *orbits, double *vr2, double *vt2, double *vp2,
                        double *vq2) {
  if(!ctx) return set_err(GALIC_ERR_ARG, "null context");
  galic::Ctx &c = ctx->c;
  galic::Particles &P = c.P;
  int n = P.n;
  HIPCHK(hipSetDevice(c.device));
  HIPCHK(hipStreamSynchronize(c.stream));
  if(vel) {
    std::vector<double> tmp(n);
    double *src[3] = {P.vx, P.vy, P.vz};
    for(int k = 0; k < 3; k++) {
      HIPCHK(hipMemcpy(tmp.data(), src[k], sizeof(double) * n, hipMemcpyDeviceToHost));
      for(int i = 0; i < n; i++) vel[3 * i + k] = tmp[i];
    }
  }
  if(orbits) HIPCHK(hipMemcpy(orbits, P.orbits, sizeof(int) * n, hipMemcpyDeviceToHost));
  if(vr2) HIPCHK(hipMemcpy(vr2, P.vr2, sizeof(double) * n, hipMemcpyDeviceToHost));
  if(vt2) HIPCHK(hipMemcpy(vt2, P.vt2, sizeof(double) * n, hipMemcpyDeviceToHost));
  if(vp2) HIPCHK(hipMemcpy(vp2, P.vp2, sizeof(double) * n, hipMemcpyDeviceToHost));
  if(vq2) HIPCHK(hipMemcpy(vq2, P.vq2, sizeof(double) * n, hipMemcpyDeviceToHost));
  return GALIC_OK;
}

int galic_calc_all_response_fields(galic_ctx *ctx) {
  if(!ctx) return set_err(GALIC_ERR_ARG, "null context");
  galic::Ctx &c = ctx->c;
  HIPCHK(hipSetDevice(c.device));
  const int n = c.P.n;
  if(n <= 0) return set_err(GALIC_ERR_STATE, "no particles set");
  for(int t = 1; t < GALIC_NTYPES; t++)
    if(type_present(c.p, t) && !c.targets_set[t]) return set_err(GALIC_ERR_STATE, "targets of type %d not set", t);
  CHK(alloc_orbit_scratch(c, n));
  HIPCHK(hipMemsetAsync(c.locA, 0, sizeof(double) * GALIC_NTYPES * galic::kFields * c.Ngrid, c.stream));
  HIPCHK(hipMemsetAsync(c.locA_d, 0, sizeof(double) * GALIC_NTYPES * galic::kFields * c.Ngrid, c.stream));
  HIPCHK(hipMemsetAsync(c.locI, 0, sizeof(double) * GALIC_NTYPES * 4 * c.Ngrid, c.stream));
  HIPCHK(hipMemsetAsync(c.locI_d, 0, sizeof(double) * GALIC_NTYPES * 4 * c.Ngrid, c.stream));
  HIPCHK(hipMemsetAsync(c.locM, 0, sizeof(double) * GALIC_NTYPES * c.Ngrid, c.stream));
  HIPCHK(hipMemsetAsync(c.fit, 0, sizeof(double) * GALIC_NTYPES * 16, c.stream));
  c.cuts_ready = false;
  CHK(build_weights(ctx, false));
  hipLaunchKernelGGL(galic::k_prepare_all, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n, 0, c.P, orbit_in(c));
  HIPCHK(hipGetLastError());
  galic::OrbitArgs A = orbit_args(c, n);
  A.loc = c.locA;
  CHK(launch_orbit64(c, A));
  hipLaunchKernelGGL(galic::k_store_orbits, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n, 0, c.o_norbit,
                     c.P.orbits);
  hipLaunchKernelGGL(galic::k_inst_init, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, c.M, c.P, c.locI, c.locM,
                     c.Ngrid);
  HIPCHK(hipGetLastError());
  CHK(global_fit(ctx));
  ctx->initial_done = true;
  return GALIC_OK;
}

int galic_rep(galic_ctx *ctx, int ntrial, const int *idx, const double *vtry, int *accept, double *dF) {
  if(!ctx || ntrial < 0 || (ntrial > 0 && (!idx || !vtry))) return set_err(GALIC_ERR_ARG, "bad rep arguments");
  galic::Ctx &c = ctx->c;
  if(!ctx->initial_done) return set_err(GALIC_ERR_STATE, "galic_calc_all_response_fields must run first");
  HIPCHK(hipSetDevice(c.device));
  for(int i = 0; i < ntrial; i++)
    if(idx[i] < 0 || idx[i] >= c.P.n) return set_err(GALIC_ERR_ARG, "trial %d: particle index %d out of range", i, idx[i]);
  const size_t slot_bytes_per_orbit = sizeof(double) * galic::kFields * c.Ngrid;
  // trials per chunk: 2 slots each, within the slot memory budget
  int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(ntrial, 1),
                                                        slot_budget(c) / (2 * slot_bytes_per_orbit)));
  CHK(alloc_trial_scratch(c, std::max(ntrial, 1)));
  CHK(alloc_orbit_scratch(c, 2 * chunk));
  CHK(alloc_slots(c, 2 * chunk, slot_bytes_per_orbit));
  if(ntrial > 0) {
    std::vector<double> tv(3 * (size_t)ntrial);
    for(int i = 0; i < ntrial; i++)
      for(int k = 0; k < 3; k++) tv[(size_t)k * ntrial + i] = vtry[3 * i + k];
    HIPCHK(hipMemcpy(c.t_idx, idx, sizeof(int) * ntrial, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.t_v, tv.data(), sizeof(double) * 3 * ntrial, hipMemcpyHostToDevice));
  }
  // init_updates (main.c:844-874): deltas are zero after every commit
  float eval_ms = 0;
  for(int base = 0; base < ntrial; base += chunk) {
    int nt = std::min(chunk, ntrial - base);
    const double *tvx = c.t_v + base, *tvy = c.t_v + ntrial + base, *tvz = c.t_v + 2 * (size_t)ntrial + base;
    hipLaunchKernelGGL(galic::k_prepare_trials, dim3(nblocks(2 * nt, 256)), dim3(256), 0, c.stream, nt,
                       c.t_idx + base, tvx, tvy, tvz, c.P, orbit_in(c));
    HIPCHK(hipGetLastError());
    galic::OrbitArgs A = orbit_args(c, 2 * nt);
    A.slots = (double *)c.slots;
    A.slot_base = 0;
    CHK(launch_orbit64(c, A));
    galic::EvalArgs E;
    memset(&E, 0, sizeof(E));
    E.P = c.P;
    E.tidx = c.t_idx + base;
    E.tvx = tvx; E.tvy = tvy; E.tvz = tvz;
    E.stacks = c.stacks;
    E.weights = c.weights;
    E.fit = c.fit;
    E.cutB_of = c.cutB_of;
    E.slots = (const double *)c.slots;
    E.slot_stride = A.slot_stride;
    E.slot_base = 0;
    E.bbox = c.o_bbox;
    E.norbit = c.o_norbit;
    E.locA_d = c.locA_d;
    E.locI_d = c.locI_d;
    E.accept = c.t_accept + base;
    E.dF = c.t_dF + base;
    for(int t = 0; t < GALIC_NTYPES; t++) {
      E.MType[t] = c.p.MType[t];
      E.fac[t][0] = c.p.fac_value_rs[t];
      E.fac[t][1] = c.p.fac_value_ts[t];
      E.fac[t][2] = c.p.fac_value_qs[t];
      E.fac[t][3] = c.p.fac_value_ps[t];
    }
    E.lowdisp = c.p.LowerDispLimit;
    // the coarse levels of the merit pyramid live in LDS up to level 7; above, each
    // workgroup of a launch of at most kCsBlocks takes a slice of HBM scratch
    const size_t ncoarse = (size_t)(c.Nstack - c.Ngrid);
    size_t lds_max = 65536;
    {
      int v = 0;
      if(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, c.device) == hipSuccess && v > 0)
        lds_max = (size_t)v;
    }
    const size_t red_bytes = sizeof(double) * 8 * 256;
    const bool cs_in_lds = sizeof(double) * ncoarse + red_bytes <= lds_max;
    const int kCsBlocks = 2048;
    E.cs_global = nullptr;
    E.cs_stride = ncoarse;
    if(!cs_in_lds) {
      if(c.cs_scratch_len < (size_t)kCsBlocks * ncoarse) {
        dfree(c.cs_scratch);
        CHK(dalloc(&c.cs_scratch, (size_t)kCsBlocks * ncoarse));
        c.cs_scratch_len = (size_t)kCsBlocks * ncoarse;
      }
      E.cs_global = c.cs_scratch;
    }
    const size_t shmem = cs_in_lds ? sizeof(double) * ncoarse + red_bytes : red_bytes;
    const int per_launch = cs_in_lds ? std::max(nt, 1) : kCsBlocks;
    HIPCHK(hipEventRecord(c.ev0, c.stream));
    for(int b0 = 0; b0 < nt; b0 += per_launch) {
      E.block_base = b0;
      hipLaunchKernelGGL(galic::k_eval_apply64, dim3(std::min(per_launch, nt - b0)), dim3(256), shmem, c.stream, c.M,
                         E);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(c.ev1, c.stream));
    HIPCHK(hipEventSynchronize(c.ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    eval_ms += ms;
  }
  c.last_eval_ms = eval_ms;
  CHK(commit(ctx));
  if(ntrial > 0) {
    if(accept) HIPCHK(hipMemcpy(accept, c.t_accept, sizeof(int) * ntrial, hipMemcpyDeviceToHost));
    if(dF) HIPCHK(hipMemcpy(dF, c.t_dF, sizeof(double) * ntrial, hipMemcpyDeviceToHost));
  }
  return GALIC_OK;
}

int galic_get_stack(galic_ctx *ctx, int type, int which, double *out) {
  if(!ctx || type < 1 || type > 3 || which < 0 || which > 15 || !out) return set_err(GALIC_ERR_ARG, "bad arguments");
  galic::Ctx &c = ctx->c;
  const int map[16] = {galic::ST_RM,  galic::ST_EMR, galic::ST_IR,  galic::ST_IT,  galic::ST_IP,  galic::ST_IQ,
                       galic::ST_RSR, galic::ST_RST, galic::ST_RSQ, galic::ST_RSP, galic::ST_MT,  galic::ST_EMT,
                       galic::ST_ER,  galic::ST_ET,  galic::ST_EP,  galic::ST_EQ};
  HIPCHK(hipSetDevice(c.device));
  HIPCHK(hipStreamSynchronize(c.stream));
  HIPCHK(hipMemcpy(out, c.stacks + ((size_t)type * galic::ST_COUNT + map[which]) * c.Nstack,
                   sizeof(double) * c.Nstack, hipMemcpyDeviceToHost));
  return GALIC_OK;
}

int galic_get_fit(galic_ctx *ctx, int type, double *out8) {
  if(!ctx || type < 1 || type > 3 || !out8) return set_err(GALIC_ERR_ARG, "bad arguments");
  galic::Ctx &c = ctx->c;
  double f[16];
  HIPCHK(hipSetDevice(c.device));
  HIPCHK(hipStreamSynchronize(c.stream));
  HIPCHK(hipMemcpy(f, c.fit + (size_t)type * 16, sizeof(f), hipMemcpyDeviceToHost));
  memcpy(out8, f, 8 * sizeof(double));
  return GALIC_OK;
}

int galic_orbit_responses(galic_ctx *ctx, int norbit, const double *pos, const double *vel, const double *mass,
                          const double *tint, const int *type, double *fields, double *t, int *orbits,
                          long long *steps, double *e0, double *e1) {
  if(!ctx || norbit < 0 || (norbit > 0 && (!pos || !vel || !mass || !tint || !type)))
    return set_err(GALIC_ERR_ARG, "bad orbit arguments");
  galic::Ctx &c = ctx->c;
  HIPCHK(hipSetDevice(c.device));
  for(int o = 0; o < norbit; o++)
    if(type[o] < 1 || type[o] > 3) return set_err(GALIC_ERR_ARG, "orbit %d: unknown type %d", o, type[o]);
  const size_t per = sizeof(double) * galic::kFields * c.Ngrid;
  int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(norbit, 1), (1ull << 30) / per));
  CHK(alloc_orbit_scratch(c, chunk));
  if(fields) CHK(alloc_slots(c, chunk, per));
  std::vector<double> tmp(chunk);
  for(int base = 0; base < norbit; base += chunk) {
    int n = std::min(chunk, norbit - base);
    double *d3[3] = {c.o_x, c.o_y, c.o_z};
    double *v3[3] = {c.o_vx, c.o_vy, c.o_vz};
    for(int k = 0; k < 3; k++) {
      for(int o = 0; o < n; o++) tmp[o] = pos[3 * (size_t)(base + o) + k];
      HIPCHK(hipMemcpy(d3[k], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
      for(int o = 0; o < n; o++) tmp[o] = vel[3 * (size_t)(base + o) + k];
      HIPCHK(hipMemcpy(v3[k], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c.o_mass, mass + base, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.o_tint, tint + base, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.o_type, type + base, sizeof(int) * n, hipMemcpyHostToDevice));
    galic::OrbitArgs A = orbit_args(c, n);
    if(fields) {
      HIPCHK(hipMemsetAsync(c.slots, 0, per * n, c.stream));
      A.slots = (double *)c.slots;
    } else {
      // integrate without binning into memory the caller wants: bin into a scratch Loc
      A.loc = c.locA_d;
    }
    CHK(launch_orbit64(c, A));
    if(!fields) HIPCHK(hipMemsetAsync(c.locA_d, 0, sizeof(double) * GALIC_NTYPES * galic::kFields * c.Ngrid, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    if(fields) {
      // slots are cell-major [cell][field]; the ABI returns 5 grids per orbit
      std::vector<double> cm((size_t)n * galic::kFields * c.Ngrid);
      HIPCHK(hipMemcpy(cm.data(), c.slots, per * n, hipMemcpyDeviceToHost));
      for(int o = 0; o < n; o++) {
        const double *src = cm.data() + (size_t)o * galic::kFields * c.Ngrid;
        double *dst = fields + (size_t)(base + o) * galic::kFields * c.Ngrid;
        for(int cell = 0; cell < c.Ngrid; cell++)
          for(int f = 0; f < galic::kFields; f++) dst[(size_t)f * c.Ngrid + cell] = src[(size_t)cell * galic::kFields + f];
      }
    }
    if(fields) HIPCHK(hipMemset(c.slots, 0, per * n));
    if(t) HIPCHK(hipMemcpy(t + base, c.o_tall, sizeof(double) * n, hipMemcpyDeviceToHost));
    if(orbits) HIPCHK(hipMemcpy(orbits + base, c.o_norbit, sizeof(int) * n, hipMemcpyDeviceToHost));
    if(steps) HIPCHK(hipMemcpy(steps + base, c.o_steps, sizeof(long long) * n, hipMemcpyDeviceToHost));
    if(e0) HIPCHK(hipMemcpy(e0 + base, c.o_e0, sizeof(double) * n, hipMemcpyDeviceToHost));
    if(e1) HIPCHK(hipMemcpy(e1 + base, c.o_e1, sizeof(double) * n, hipMemcpyDeviceToHost));
  }
  return GALIC_OK;
}

int galic_fp32_orbit_responses(galic_ctx *ctx, int norbit, const double *pos, const double *vel, const double *mass,
                               const double *tint, const int *type, double *fields, double *t, int *orbits,
                               long long *steps, double *e0, double *e1) {
  if(!ctx || norbit < 0 || (norbit > 0 && (!pos || !vel || !mass || !tint || !type)))
    return set_err(GALIC_ERR_ARG, "bad orbit arguments");
  galic::Ctx &c = ctx->c;
  HIPCHK(hipSetDevice(c.device));
  for(int o = 0; o < norbit; o++)
    if(type[o] < 1 || type[o] > 3) return set_err(GALIC_ERR_ARG, "orbit %d: unknown type %d", o, type[o]);
  const size_t per = sizeof(float) * galic::kFields * c.Ngrid;
  int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(norbit, 1), (1ull << 30) / per));
  CHK(alloc_orbit_scratch(c, chunk));
  CHK(alloc_slots(c, chunk, per));
  float *fin = nullptr;
  HIPCHK(hipMalloc(&fin, sizeof(float) * 6 * chunk));
  std::vector<double> tmp(chunk);
  std::vector<float> cm(fields ? (size_t)chunk * galic::kFields * c.Ngrid : 0);
  int rc = GALIC_OK;
  for(int base = 0; base < norbit && rc == GALIC_OK; base += chunk) {
    const int n = std::min(chunk, norbit - base);
    double *d3[3] = {c.o_x, c.o_y, c.o_z};
    double *v3[3] = {c.o_vx, c.o_vy, c.o_vz};
    for(int k = 0; k < 3 && rc == GALIC_OK; k++) {
      for(int o = 0; o < n; o++) tmp[o] = pos[3 * (size_t)(base + o) + k];
      rc = hip_ok(hipMemcpy(d3[k], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
      for(int o = 0; o < n; o++) tmp[o] = vel[3 * (size_t)(base + o) + k];
      if(rc == GALIC_OK) rc = hip_ok(hipMemcpy(v3[k], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    if(rc == GALIC_OK) rc = hip_ok(hipMemcpy(c.o_mass, mass + base, sizeof(double) * n, hipMemcpyHostToDevice));
    if(rc == GALIC_OK) rc = hip_ok(hipMemcpy(c.o_tint, tint + base, sizeof(double) * n, hipMemcpyHostToDevice));
    if(rc == GALIC_OK) rc = hip_ok(hipMemcpy(c.o_type, type + base, sizeof(int) * n, hipMemcpyHostToDevice));
    if(rc == GALIC_OK) rc = hip_ok(hipMemsetAsync(c.slots, 0, per * n, c.stream));
    if(rc != GALIC_OK) break;
    galic::OrbitArgs32 A = orbit_args32(c, n);
    A.slots = (float *)c.slots;
    A.fin = fin;
    rc = launch_orbit32(c, A);
    if(rc != GALIC_OK) break;
    if(e0 || e1) {
      hipLaunchKernelGGL(galic::k_energy32, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, c.M, n,
                         (const double *)c.o_x, (const double *)c.o_y, (const double *)c.o_z, (const double *)c.o_vx,
                         (const double *)c.o_vy, (const double *)c.o_vz, (const float *)fin, c.o_e0, c.o_e1);
      rc = hip_ok(hipGetLastError());
    }
    if(rc == GALIC_OK) rc = hip_ok(hipStreamSynchronize(c.stream));
    if(rc == GALIC_OK && fields) {
      rc = hip_ok(hipMemcpy(cm.data(), c.slots, per * n, hipMemcpyDeviceToHost));
      for(int o = 0; o < n && rc == GALIC_OK; o++) {
        const float *src = cm.data() + (size_t)o * galic::kFields * c.Ngrid;
        double *dst = fields + (size_t)(base + o) * galic::kFields * c.Ngrid;
        for(int cell = 0; cell < c.Ngrid; cell++)
          for(int f = 0; f < galic::kFields; f++) dst[(size_t)f * c.Ngrid + cell] = src[(size_t)cell * galic::kFields + f];
      }
    }
    if(rc == GALIC_OK) rc = hip_ok(hipMemset(c.slots, 0, per * n));
    if(rc == GALIC_OK && t) rc = hip_ok(hipMemcpy(t + base, c.o_tall, sizeof(double) * n, hipMemcpyDeviceToHost));
    if(rc == GALIC_OK && orbits) rc = hip_ok(hipMemcpy(orbits + base, c.o_norbit, sizeof(int) * n, hipMemcpyDeviceToHost));
    if(rc == GALIC_OK && steps) rc = hip_ok(hipMemcpy(steps + base, c.o_steps, sizeof(long long) * n, hipMemcpyDeviceToHost));
    if(rc == GALIC_OK && e0) rc = hip_ok(hipMemcpy(e0 + base, c.o_e0, sizeof(double) * n, hipMemcpyDeviceToHost));
    if(rc == GALIC_OK && e1) rc = hip_ok(hipMemcpy(e1 + base, c.o_e1, sizeof(double) * n, hipMemcpyDeviceToHost));
  }
  (void)hipFree(fin);
  return rc;
}

int galic_set_dispersions(galic_ctx *ctx, int n, const double *disp4) {
  if(!ctx || !disp4) return set_err(GALIC_ERR_ARG, "null argument");
  galic::Ctx &c = ctx->c;
  if(n != c.P.n) return set_err(GALIC_ERR_ARG, "dispersions for %d particles, model has %d", n, c.P.n);
  HIPCHK(hipSetDevice(c.device));
  dfree(c.P.disp);
  CHK(dalloc(&c.P.disp, 4 * (size_t)n));
  HIPCHK(hipMemcpy(c.P.disp, disp4, sizeof(double) * 4 * n, hipMemcpyHostToDevice));
  return GALIC_OK;
}

int galic_fp32_calc_all_response_fields(galic_ctx *ctx) {
  if(!ctx) return set_err(GALIC_ERR_ARG, "null context");
  galic::Ctx &c = ctx->c;
  HIPCHK(hipSetDevice(c.device));
  const int n = c.P.n;
  if(n <= 0) return set_err(GALIC_ERR_STATE, "no particles set");
  for(int t = 1; t < GALIC_NTYPES; t++)
    if(type_present(c.p, t) && !c.targets_set[t]) return set_err(GALIC_ERR_STATE, "targets of type %d not set", t);
  CHK(alloc_orbit_scratch(c, n));
  const size_t na = (size_t)GALIC_NTYPES * galic::kFields * c.Ngrid, ni = (size_t)GALIC_NTYPES * 4 * c.Ngrid,
               nm = (size_t)GALIC_NTYPES * c.Ngrid;
  HIPCHK(hipMemsetAsync(c.locA_q, 0, sizeof(long long) * na, c.stream));
  HIPCHK(hipMemsetAsync(c.locA_dq, 0, sizeof(long long) * na, c.stream));
  HIPCHK(hipMemsetAsync(c.locI_q, 0, sizeof(long long) * ni, c.stream));
  HIPCHK(hipMemsetAsync(c.locI_dq, 0, sizeof(long long) * ni, c.stream));
  HIPCHK(hipMemsetAsync(c.locM_q, 0, sizeof(long long) * nm, c.stream));
  HIPCHK(hipMemsetAsync(c.fit, 0, sizeof(double) * GALIC_NTYPES * 16, c.stream));
  c.cuts_ready = false;
  CHK(build_weights(ctx, false));
  hipLaunchKernelGGL(galic::k_prepare_all, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n, 0, c.P, orbit_in(c));
  HIPCHK(hipGetLastError());
  // the orbits into response logs, chunk by chunk as the arena holds them, each chunk's
  // records then added to the fixed-point Loc grids (the logging kernel keeps its lanes
  // integrating; atomics at every flush would stall them)
  log_setup(c);
  CHK(fp32_initial_logs(c, n));
  hipLaunchKernelGGL(galic::k_store_orbits, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n, 0, c.o_norbit,
                     c.P.orbits);
  hipLaunchKernelGGL(galic::k_store_steps, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n,
                     (const long long *)c.o_steps, c.psteps);
  hipLaunchKernelGGL(galic::k_inst_init_q, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, c.M, c.P, c.locI_q,
                     c.locM_q, c.Ngrid, c.fp32_iscale, c.fp32_scale[0]);
  HIPCHK(hipGetLastError());
  CHK(commit_q(c));
  // EG_MassLoc in FP64 from its fixed-point sum (the zero delta leaves it unchanged)
  hipLaunchKernelGGL(galic::k_commit_q, dim3(nblocks(nm, 256)), dim3(256), 0, c.stream, c.locM_q, c.locA_dq, c.locM,
                     (size_t)c.Ngrid, 1, (const double *)c.qinv, nm);
  HIPCHK(hipGetLastError());
  CHK(global_fit(ctx));
  ctx->fp32_initial_done = true;
  ctx->initial_done = false;
  return GALIC_OK;
}

int galic_fp32_autotune(galic_ctx *ctx, int nsample, int *variant, int *blocks_per_cu) {
  if(!ctx || nsample < 1) return set_err(GALIC_ERR_ARG, "bad autotune arguments");
  galic::Ctx &c = ctx->c;
  if(!ctx->fp32_initial_done) return set_err(GALIC_ERR_STATE, "galic_fp32_calc_all_response_fields must run first");
  HIPCHK(hipSetDevice(c.device));
  // no more orbits than the response logs of one launch fit in half the memory the arena may
  // take (the other half covers the blocks the lanes claim and leave unused)
  const double arena_blocks = (double)slot_budget(c) / log_block_bytes();
  const double per_orbit = (c.log_est_blocks + 1.0) * 1.25;
  const int fit_n = (int)std::max(1024.0, std::min(2e9, 0.5 * arena_blocks / per_orbit));
  const int n = std::min(std::min(nsample, c.P.n), fit_n);
  CHK(alloc_orbit_scratch(c, n));
  // a strided sample of the model's own orbits at their current velocities,
  // longest expected first as in a rep
  std::vector<int> idx(n);
  for(int i = 0; i < n; i++) idx[i] = (int)((long long)i * c.P.n / n);
  CHK(alloc_trial_scratch(c, n));
  HIPCHK(hipMemcpy(c.t_idx, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  if(galic::sort_trials_by_steps(n, c.t_idx, c.psteps, c.t_key, c.t_key2, c.t_ord, c.t_ord2, &c.sort_tmp,
                                 &c.sort_tmp_bytes, c.stream) != 0)
    return set_err(GALIC_ERR_HIP, "sorting the tuning sample failed");
  hipLaunchKernelGGL(galic::k_prepare_ktrials, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n, 0,
                     (const int *)c.t_ord2, c.t_idx, (const double *)nullptr, (const double *)nullptr,
                     (const double *)nullptr, c.P, orbit_in(c));
  HIPCHK(hipGetLastError());
  // grid sizes in workgroups per CU (the results are bit-identical; only the schedule differs)
  int occ = 4;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void *>(galic::k_orbit32_log), 256, 0));
  occ = std::max(1, std::min(occ, 16));
  // every count from 1 to the resident limit, and twice that (a persistent grid larger than
  // the CUs hold only queues)
  const int variants[1] = {0};
  std::vector<int> bpcs;
  for(int b = 1; b <= occ; b++) bpcs.push_back(b);
  bpcs.push_back(2 * occ);
  double best = 1e300;
  int bv = 0, bb = 4;
  const long long steps0 = c.acc_steps, orbits0 = c.acc_orbits;
  const double ms0 = c.acc_orbit_ms;
  const long long launches0 = c.acc_orbit_launches;
  for(int v : variants)
    for(int b : bpcs) {
      c.fp32_variant = v;
      c.fp32_bpc = b;
      long long s_before = c.acc_steps;
      int lr = kLogArenaFull;
      for(int attempt = 0; attempt < 6 && lr == kLogArenaFull; attempt++) {
        const size_t bb = log_block_bytes();
        CHK(ensure_arena(c, (size_t)(log_blocks_planned(c, n) * bb), (size_t)(log_blocks_planned(c, n) * bb)));
        lr = launch_orbit32_log(c, orbit_args32(c, n));
      }
      if(lr == kLogArenaFull) return set_err(GALIC_ERR_NOMEM, "no room for the tuning sample's response logs");
      CHK(lr);
      double t = c.last_orbit_ms / std::max<long long>(1, c.acc_steps - s_before);
      if(t < best) {
        best = t;
        bv = v;
        bb = b;
      }
    }
  // tuning launches are not part of the caller's measurements
  c.acc_steps = steps0;
  c.acc_orbits = orbits0;
  c.acc_orbit_ms = ms0;
  c.acc_orbit_launches = launches0;
  c.fp32_variant = bv;
  c.fp32_bpc = bb;
  c.fp32_tuned = true;
  HIPCHK(hipStreamSynchronize(c.stream));
  if(variant) *variant = bv;
  if(blocks_per_cu) *blocks_per_cu = bb;
  return GALIC_OK;
}

int galic_fp32_set_launch(galic_ctx *ctx, int variant, int blocks_per_cu) {
  if(!ctx || variant != 0 || blocks_per_cu < 1 || blocks_per_cu > 32)
    return set_err(GALIC_ERR_ARG, "bad FP32 launch (variant 0, 1..32 workgroups per CU)");
  ctx->c.fp32_variant = variant;
  ctx->c.fp32_bpc = blocks_per_cu;
  ctx->c.fp32_tuned = true;
  return GALIC_OK;
}

namespace {

// Reps of the batched FP32 path.  ntrial trials (K per particle idx[i]) are
// taken in reps of rep_size: every trial of a rep is evaluated against the grids
// as they were at the start of the rep, and the rep's deltas are committed
// before the next rep (main.c:93-100).  An orbit depends only on its particle's
// position and velocity, and particles are distinct within a group of reps, so
// the orbits of as many whole reps as the slot budget holds are integrated in one
// launch, longest expected orbit first; the reps are then evaluated and
// committed in turn.  The result is that of one rep at a time.
int fp32_run(galic_ctx *ctx, int ntrial, const int *idx, int K, int rep_size, const double *vtry, double sigma_frac,
             unsigned long long seed, int *accept) {
  if(!ctx || ntrial < 0 || (ntrial > 0 && !idx) || K < 1 || K > 64 || rep_size < 1)
    return set_err(GALIC_ERR_ARG, "bad FP32 rep arguments (K must be 1..64, rep_size >= 1)");
  galic::Ctx &c = ctx->c;
  if(!ctx->fp32_initial_done) return set_err(GALIC_ERR_STATE, "galic_fp32_calc_all_response_fields must run first");
  if(!vtry && !c.P.disp) return set_err(GALIC_ERR_STATE, "device trial draws need galic_set_dispersions");
  HIPCHK(hipSetDevice(c.device));
  for(int i = 0; i < ntrial; i++)
    if(idx[i] < 0 || idx[i] >= c.P.n) return set_err(GALIC_ERR_ARG, "trial %d: particle index %d out of range", i, idx[i]);
  CHK(alloc_trial_scratch(c, std::max(ntrial, 1)));
  size_t nk = (size_t)std::max(ntrial, 1) * K;
  if(nk > c.max_ktrials) {
    dfree(c.t_vk);
    dfree(c.t_dFk);
    dfree(c.t_vinst);
    CHK(dalloc(&c.t_vk, 3 * nk));
    CHK(dalloc(&c.t_dFk, nk));
    CHK(dalloc(&c.t_vinst, nk));
    c.max_ktrials = nk;
  }
  c.last_ntrial = ntrial;
  c.last_k = K;
  double *vtx = c.t_vk, *vty = c.t_vk + nk, *vtz = c.t_vk + 2 * nk;
  if(ntrial > 0) {
    HIPCHK(hipMemcpy(c.t_idx, idx, sizeof(int) * ntrial, hipMemcpyHostToDevice));
    if(vtry) {
      std::vector<double> tv(3 * nk);
      for(size_t q = 0; q < (size_t)ntrial * K; q++)
        for(int k = 0; k < 3; k++) tv[k * nk + q] = vtry[3 * q + k];
      HIPCHK(hipMemcpy(c.t_vk, tv.data(), sizeof(double) * 3 * nk, hipMemcpyHostToDevice));
    }
  }
  // the log arena: as much of the free memory as the call could use, kept across calls; split
  // between the two orbit sets when the call's orbits need more than one launch, so the next
  // group's orbits integrate while a group is judged
  bool overlap = false;
  {
    const size_t bb = log_block_bytes();
    const size_t all = (size_t)(log_blocks_planned(c, std::max(ntrial, 1) * (K + 1)) * bb);
    const size_t one = (size_t)(log_blocks_planned(c, std::min(std::max(ntrial, 1), 1024) * (K + 1)) * bb);
    const size_t budget = slot_budget(c);
    overlap = c.overlap == 1 || (c.overlap < 0 && all > budget / 2);
    if(overlap) {
      CHK(ensure_alt_set(c));
      // each set's arena about half of what one arena could take (the one left by earlier calls
      // shrinks first, so that the other has room)
      const size_t half = std::max(one, (budget + c.alt.slot_bytes) / 2);
      auto fit_half = [&]() -> int {
        if(c.slot_bytes >= one && c.slot_bytes <= half + half / 8 && 2 * c.slot_bytes >= half) return GALIC_OK;
        if(c.slots) (void)hipFree(c.slots);
        c.slots = nullptr;
        c.slot_bytes = 0;
        return ensure_arena(c, half, one);
      };
      CHK(fit_half());
      swap_orbit_set(c);
      const int rc = fit_half();
      swap_orbit_set(c);
      CHK(rc);
    } else {
      CHK(ensure_arena(c, std::min(all, budget), one));
    }
  }
  // the orbit launches beside a group's judging share the CUs with the merit kernel: their own
  // workgroups per CU (set around those launches; restored on every return)
  struct Restore {
    int &ref;
    int saved;
    ~Restore() { ref = saved; }
  } bpc_restore{c.fp32_bpc, c.fp32_bpc};
  const int bpc_alone = c.fp32_bpc, bpc_beside = c.fp32_bpc_overlap > 0 ? c.fp32_bpc_overlap : c.fp32_bpc;
  // trials whose orbits' logs fit in the arena at the current estimate (and the caller's cap)
  auto group_cap = [&]() {
    const double blocks = (double)(c.slot_bytes / log_block_bytes());
    const double per_trial = (K + 1) * (c.log_est_blocks + 1.0) * 1.25;
    const double lanes = (double)orbit_grid(std::max(ntrial, 1) * (K + 1), 256, c.fp32_bpc) * 256;
    double cap = std::min((double)std::max(ntrial, 1), (blocks - 64 - lanes * galic::LOG_CHUNK) / per_trial);
    if(c.max_group_trials > 0) cap = std::min(cap, (double)c.max_group_trials);
    return (int)std::max(1.0, cap);
  };
  int cap = group_cap();
  const int nrep = (ntrial + rep_size - 1) / rep_size;
  // rep r covers trials [r*rep_size, min((r+1)*rep_size, ntrial)); a group is a run of
  // whole reps whose particles are distinct and whose orbits fit the slots (or a
  // part of one rep when a single rep does not fit)
  std::vector<char> seen(c.P.n, 0);
  std::vector<int> touched;
  const size_t nbw = (size_t)GALIC_NTYPES * galic::kFields * c.Nstack;
  // all five fields in one pass when their fine and coarse grids fit in the LDS of one workgroup
  size_t lds_max = 65536;
  {
    int v = 0;
    if(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, c.device) == hipSuccess && v > 0)
      lds_max = (size_t)v;
  }
  // the merit kernel's LDS: per-wave partials, block sums, the coarse nodes' terms, one or two
  // int64 grids and the K+1 block tables; two grids where they fit
  auto eval_lds = [&](int ngrids) {
    return (galic::MERIT_MAX_WAVES * (size_t)K + 64) * sizeof(double) +
           sizeof(double2) * (size_t)GALIC_NSTACK(std::max(c.L - 2, 0)) +
           ngrids * sizeof(long long) * (size_t)c.Nbin * (c.Nbin + (ngrids == 2 ? galic::MERIT_ROW_PAD : 0)) +
           sizeof(int) * (size_t)(K + 1) * c.log_M + 1024;
  };
  // The merit kernel's form.  Tiles (k_eval_tile32): the grid in LDS tiles of 32 x 32 cells,
  // the default above 128^2 cells and wherever the LDS grids do not fit.  Otherwise one or two
  // LDS grids of the whole response grid (k_eval_log32).  The HBM-scratch form of k_eval_log32
  // (two int64 grids per workgroup in HBM) is kept for grids above 128^2 when tiles are
  // switched off or K exceeds what the tile lists hold.
  // with the tile's (residual, weight) pairs in LDS where they fit (K up to about 8 on 512^2 grids)
  auto tile_lds_bw = [&](bool bwl) { return galic::tile_lds_bytes(c.L, K, c.log_M, bwl) + 1024; };
  auto tile_lds = [&]() { return tile_lds_bw(false); };
  const bool tiles_ok = c.L >= 5 && K <= galic::TILE_MAX_K;
  // (the block tables in LDS when they leave room, else read from memory)
  auto sparse_tbl = [&]() { return galic::sparse_lds_bytes(c.L, K, c.log_M, true) + 1024 <= lds_max; };
  auto sparse_lds = [&]() { return galic::sparse_lds_bytes(c.L, K, c.log_M, sparse_tbl()) + 1024; };
  auto use_tiles = [&]() {
    if(!tiles_ok || c.merit_form == 2) return false;
    if(c.merit_form == 1 || c.merit_form == 3 || c.L > 7) return tile_lds() <= lds_max;
    return eval_lds(1) > lds_max && tile_lds() <= lds_max;
  };
  // tiles restricted to the nodes the logs touch, where their LDS fits (form 3; the dense
  // tiles measured faster so far and are the automatic choice)
  auto use_sparse = [&]() { return use_tiles() && c.merit_form == 3 && sparse_lds() <= lds_max; };
  // 16x16 tiles in the particle's own int32 fixed point (k_eval_t16): the automatic choice
  // where it applies (form 4 asks for it)
  const size_t t16_lds = galic::t16_lds_bytes(std::max(c.L, 4), K, K <= 8 ? 8 : 16) + 1024;
  const bool t16 = (c.merit_form == 0 || c.merit_form == 4) && c.L >= 4 && c.L <= galic::T16_MAX_LEVEL &&
                   K <= galic::T16_MAX_K && t16_lds <= lds_max;
  const bool gmem = c.L > 7;
  const int kGBlocks = 512;
  auto eval_lds_gmem = [&]() {
    return galic::MERIT_MAX_WAVES * (size_t)K * sizeof(double) + sizeof(long long) * (size_t)GALIC_NSTACK(c.L - 4) +
           2 * sizeof(unsigned int) * (((size_t)c.Ngrid / 4 + 31) / 32) + sizeof(int) * (size_t)(K + 1) * c.log_M +
           1024;
  };
  if(gmem && !t16 && !use_tiles()) {
    const size_t need = (size_t)kGBlocks * 2 * c.Ngrid;
    if(c.g32_scratch_len < need) {
      dfree(c.g32_scratch);
      CHK(dalloc(&c.g32_scratch, need));  // zero; the kernel leaves it zero
      c.g32_scratch_len = need;
    }
  }
  bool packed = false;
  float eval_ms = 0;
  // the end of the group that starts at trial s0
  auto plan = [&](int s0) -> int {
    int s1;
    const int r0 = s0 / rep_size;
    const int rep_end0 = std::min(ntrial, (r0 + 1) * rep_size);
    if(rep_end0 - s0 > cap || s0 % rep_size != 0) {
      s1 = std::min(rep_end0, s0 + cap);  // (part of) one rep
    } else {
      s1 = rep_end0;
      for(int t : touched) seen[t] = 0;
      touched.clear();
      for(int i = s0; i < s1; i++)
        if(!seen[idx[i]]) {
          seen[idx[i]] = 1;
          touched.push_back(idx[i]);
        }
      while(s1 < ntrial) {
        const int e = std::min(ntrial, s1 + rep_size);
        if(e - s0 > cap) break;
        bool distinct = true;
        for(int i = s1; i < e && distinct; i++) distinct = !seen[idx[i]];
        if(!distinct) break;
        for(int i = s1; i < e; i++)
          if(!seen[idx[i]]) {
            seen[idx[i]] = 1;
            touched.push_back(idx[i]);
          }
        s1 = e;
      }
    }
    return s1;
  };
  // trial draws, orbit order and the orbits' launch of group [s0, s1) on the active orbit set
  auto start = [&](int s0, int s1, galic::OrbitArgs32 &A) -> int {
    const int ng = s1 - s0;
    CHK(alloc_sort_scratch(c, std::max(ntrial, 1)));
    if(!vtry) {
      // trials drawn from the particles' current velocities (unchanged by earlier reps
      // of the group: its particles are distinct)
      hipLaunchKernelGGL(galic::k_draw_trials, dim3(nblocks((size_t)ng * K, 256)), dim3(256), 0, c.stream, c.M, ng, K,
                         std::min(c.fp32_nfresh, K), c.t_idx + s0, c.P, (const double *)c.P.disp, sigma_frac,
                         c.p.MaxVelInUnitsVesc, (uint32_t)seed, (uint32_t)(3000 + (seed >> 32)),
                         (const int *)c.tovs_dev, vtx + (size_t)s0 * K, vty + (size_t)s0 * K, vtz + (size_t)s0 * K);
      HIPCHK(hipGetLastError());
    }
    if(galic::sort_trials_by_steps(ng, c.t_idx + s0, c.psteps, c.t_key, c.t_key2, c.t_ord, c.t_ord2, &c.sort_tmp,
                                   &c.sort_tmp_bytes, c.stream) != 0)
      return set_err(GALIC_ERR_HIP, "sorting the trials by expected orbit length failed");
    hipLaunchKernelGGL(galic::k_invert_order, dim3(nblocks(ng, 256)), dim3(256), 0, c.stream, ng,
                       (const int *)c.t_ord2, c.t_inv);
    CHK(alloc_orbit_scratch(c, ng * (K + 1)));
    hipLaunchKernelGGL(galic::k_prepare_ktrials, dim3(nblocks((size_t)ng * (K + 1), 256)), dim3(256), 0, c.stream, ng,
                       K, (const int *)c.t_ord2, (const int *)(c.t_idx + s0), (const double *)(vtx + (size_t)s0 * K),
                       (const double *)(vty + (size_t)s0 * K), (const double *)(vtz + (size_t)s0 * K), c.P,
                       orbit_in(c));
    HIPCHK(hipGetLastError());
    A = orbit_args32(c, ng * (K + 1));
    return orbit32_log_begin(c, A);
  };
  // waits for the active set's launch: a table that overflowed grows and the launch repeats; if
  // the arena ran out, the group is planned again, smaller, and launched again
  auto finish = [&](int s0, int &s1, galic::OrbitArgs32 &A) -> int {
    for(int attempt = 0; attempt < 64; attempt++) {
      const int r = orbit32_log_end(c, A);
      if(r == GALIC_OK) return GALIC_OK;
      if(r == kLogTableFull) {
        CHK(orbit32_log_begin(c, A));
        continue;
      }
      if(r != kLogArenaFull) return r;
      if(s1 - s0 <= 1) return set_err(GALIC_ERR_NOMEM, "the response logs of one trial do not fit in device memory");
      cap = group_cap();
      s1 = plan(s0);
      CHK(start(s0, s1, A));
    }
    return set_err(GALIC_ERR_NOMEM, "orbit response logs still overflow after growing the block tables");
  };
  // the reps of group [s0, s1), its orbits integrated: merit, choice, commit and fit per rep
  auto judge = [&](int s0, int s1) -> int {
    const int r0 = s0 / rep_size;
    const int ng = s1 - s0;
    // (the block tables may have grown with this launch: the form is chosen after it)
    const bool tiles = !t16 && use_tiles();
    const bool sparse = !t16 && use_sparse();
    // without tiles, the HBM-scratch form above 128^2 cells and wherever one LDS grid does not fit
    const bool gmem_now = !t16 && !tiles && (gmem || eval_lds(1) > lds_max);
    const bool use_o = gmem_now || eval_lds(2) <= lds_max;
    const bool bwl = tiles && tile_lds_bw(true) <= lds_max;
    const size_t shmem = (t16      ? t16_lds
                          : sparse ? sparse_lds()
                          : tiles  ? tile_lds_bw(bwl)
                          : gmem_now ? eval_lds_gmem()
                                     : eval_lds(use_o ? 2 : 1)) -
                         1024;
    if(shmem + 1024 > lds_max)
      return set_err(GALIC_ERR_ARG, "the merit kernel needs %zu bytes of LDS, the device gives %zu", shmem, lds_max);
    if(gmem_now && c.g32_scratch_len < (size_t)kGBlocks * 2 * c.Ngrid) {
      dfree(c.g32_scratch);
      CHK(dalloc(&c.g32_scratch, (size_t)kGBlocks * 2 * c.Ngrid));
      c.g32_scratch_len = (size_t)kGBlocks * 2 * c.Ngrid;
    }
    // tile form: workgroups resident at once (each takes trials b, b + grid, ...), and list
    // room for the largest particle of the group (a record is listed in at most four tiles)
    int tile_grid = 0, tile_cap = 0;
    const int t16_kr = K <= 8 ? 8 : 16;
    const void *t16_kern =
        c.L == 9 ? (t16_kr == 8 ? (const void *)galic::k_eval_t16<9, 8> : (const void *)galic::k_eval_t16<9, 16>)
                 : (t16_kr == 8 ? (const void *)galic::k_eval_t16<0, 8> : (const void *)galic::k_eval_t16<0, 16>);
    if(tiles || t16) {
      int per_cu = 0, ncu = 0;
      auto kern = t16 ? t16_kern
                  : sparse ? (c.L == 9 ? (const void *)galic::k_eval_sparse32<9> : (const void *)galic::k_eval_sparse32<0>)
                  : c.L == 9 ? (bwl ? (const void *)galic::k_eval_tile32<9, true> : (const void *)galic::k_eval_tile32<9, false>)
                             : (bwl ? (const void *)galic::k_eval_tile32<0, true> : (const void *)galic::k_eval_tile32<0, false>);
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, t16 ? 64 * galic::kFields : 1024, shmem));
      HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device));
      if(per_cu < 1) return set_err(GALIC_ERR_ARG, "the tile merit kernel does not fit on a CU (%zu bytes of LDS)", shmem);
      if(c.merit_wg_per_cu > 0) per_cu = std::min(per_cu, c.merit_wg_per_cu);
      tile_grid = std::min(ng, per_cu * ncu);
      if(getenv("GALIC_VERBOSE")) {
        hipFuncAttributes fa;
        if(hipFuncGetAttributes(&fa, kern) == hipSuccess)
          fprintf(stderr, "[galic] merit kernel: %d workgroups per CU x %d CUs, %zu B dynamic LDS, %d VGPRs, "
                          "%zu B static LDS, %zu B scratch per lane\n",
                  per_cu, ncu, shmem, fa.numRegs, fa.sharedSizeBytes, fa.localSizeBytes);
      }
      if(!c.tile_aux) CHK(dalloc(&c.tile_aux, 4));
      HIPCHK(hipMemsetAsync(c.tile_aux, 0, 2 * sizeof(unsigned int), c.stream));
      hipLaunchKernelGGL(galic::k_log_totals, dim3(nblocks(ng, 256)), dim3(256), 0, c.stream, ng, K,
                         (const int *)nullptr, (const int *)c.log_n, c.tile_aux);
      HIPCHK(hipGetLastError());
      unsigned int mx = 0;
      HIPCHK(hipMemcpyAsync(&mx, c.tile_aux, sizeof(mx), hipMemcpyDeviceToHost, c.stream));
      HIPCHK(hipStreamSynchronize(c.stream));
      const size_t want = 4 * (size_t)mx + 64;
      if(want > 0x7fffffffu) return set_err(GALIC_ERR_NOMEM, "a particle's response logs hold %u records", mx);
      tile_cap = (int)((want + 3) & ~(size_t)3);
      // the sparse form keeps a copy of each listed record (a word and five float4, 21 words);
      // the dense form one word per entry, the 16x16 form two (the arena record index and the log)
      if(t16 && (size_t)c.slot_bytes / log_block_bytes() * galic::LOG_B > 0xffffffffull)
        return set_err(GALIC_ERR_ARG, "the log arena has more records than 32-bit list entries address");
      const size_t words = (size_t)tile_grid * tile_cap * (t16 ? 2 : sparse ? 21 : 1);
      if(c.tile_list_len < words) {
        dfree(c.tile_list);
        c.tile_list_len = 0;
        CHK(dalloc(&c.tile_list, words));
        c.tile_list_len = words;
      }
    }
    // the merit launch arguments of trials [a, ...) of the group
    auto eval_args = [&](int a) {
      galic::EvalArgs32 E;
      memset(&E, 0, sizeof(E));
      E.P = c.P;
      E.K = K;
      E.trial_base = a;
      E.slot_pos = c.t_inv + (a - s0);
      E.tidx = c.t_idx;
      E.tvx = vtx;
      E.tvy = vty;
      E.tvz = vtz;
      E.stacks = c.stacks;
      E.merit_bw = c.merit_bw;
      E.merit_bwf = c.merit_bwf;
      E.fit = c.fit;
      E.cutB_of = c.cutB_of;
      E.log_blocks = c.log_blocks;
      E.log_n = c.log_n;
      E.log_M = c.log_M;
      E.log_q = c.log_q;
      E.log_v = c.log_v;
      for(int f = 0; f < galic::kFields; f++) {
        E.log_scale[f] = c.log_scale[f];
        E.log_unit[f] = c.log_unit[f];
        E.log_to_loc[f] = c.log_to_loc[f];
      }
      E.bbox = c.o_bbox;
      E.norbit = c.o_norbit;
      E.osteps = c.o_steps;
      E.psteps = c.psteps;
      E.locA_dq = c.locA_dq;
      E.locI_dq = c.locI_dq;
      E.accept = c.t_accept;
      E.dF = c.t_dF;
      E.dFk = c.t_dFk;
      for(int t = 0; t < GALIC_NTYPES; t++) {
        E.MType[t] = c.p.MType[t];
        E.fac[t][0] = c.p.fac_value_rs[t];
        E.fac[t][1] = c.p.fac_value_ts[t];
        E.fac[t][2] = c.p.fac_value_qs[t];
        E.fac[t][3] = c.p.fac_value_ps[t];
      }
      E.lowdisp = c.p.LowerDispLimit;
      for(int f = 0; f < 5; f++) E.qscale[f] = c.fp32_scale[f];
      E.iscale = c.fp32_iscale;
      E.gscratch = c.g32_scratch;
      E.no_deltas = 1;
      return E;
    };
    // the 16x16-tile merit of trials [a, b) of the group and the write-back of each choice; the
    // launch order by decreasing record count made by rocPRIM's radix sort, or in one workgroup
    // (device_order, which a graph can hold: the radix sort sets memory with memset nodes)
    auto t16_rep = [&](galic::EvalArgs32 &E, int a, int b, bool device_order) -> int {
      E.nblk = b - a;
      E.tlist2 = (uint2 *)c.tile_list;
      E.tlist_cap = tile_cap;
      E.tile_err = c.tile_aux + 1;
      E.log_tot = c.log_tot;
      E.vinst = c.t_vinst;
      E.prof = c.merit_prof;
      E.tcut = c.tcut;
      E.work_order = c.t_ord2;
      E.work_next = (int *)c.tile_aux + 2;
      if(device_order) {
        int m = 1;
        while(m < b - a) m <<= 1;
        hipLaunchKernelGGL(galic::k_record_order, dim3(1), dim3(1024), sizeof(unsigned long long) * m, c.stream, b - a,
                           K, E.slot_pos, (const int *)c.log_n, c.t_ord2);
        hipLaunchKernelGGL(galic::k_fill_u32, dim3(1), dim3(64), 0, c.stream, c.tile_aux + 2, 0u, (size_t)1);
      } else {
        if(galic::sort_trials_by_records(b - a, K, E.slot_pos, c.log_n, c.t_key, c.t_key2, c.t_ord, c.t_ord2,
                                         &c.sort_tmp, &c.sort_tmp_bytes, c.stream) != 0)
          return set_err(GALIC_ERR_HIP, "sorting the trials by record count failed");
        HIPCHK(hipMemsetAsync(E.work_next, 0, sizeof(int), c.stream));
      }
      const int grid = std::min(b - a, tile_grid);
      // instantaneous terms before, the write-back of the choice after the merit kernel
      hipLaunchKernelGGL(galic::k_trial_inst, dim3(nblocks((size_t)(b - a) * K, 256)), dim3(256), 0, c.stream, c.M, E,
                         c.t_vinst);
      const dim3 tb(64 * galic::kFields);
      if(c.L == 9 && t16_kr == 8)
        hipLaunchKernelGGL((galic::k_eval_t16<9, 8>), dim3(grid), tb, shmem, c.stream, c.M, E);
      else if(c.L == 9)
        hipLaunchKernelGGL((galic::k_eval_t16<9, 16>), dim3(grid), tb, shmem, c.stream, c.M, E);
      else if(t16_kr == 8)
        hipLaunchKernelGGL((galic::k_eval_t16<0, 8>), dim3(grid), tb, shmem, c.stream, c.M, E);
      else
        hipLaunchKernelGGL((galic::k_eval_t16<0, 16>), dim3(grid), tb, shmem, c.stream, c.M, E);
      HIPCHK(hipGetLastError());
      hipLaunchKernelGGL(galic::k_apply_best, dim3(nblocks(b - a, 256)), dim3(256), 0, c.stream, c.M, E);
      HIPCHK(hipGetLastError());
      return GALIC_OK;
    };
    // the commit arguments of trials [a, b): the taken trials' orbits into the Loc deltas, the
    // orbits they replace out
    auto commit_args = [&](const galic::EvalArgs32 &E, int a, int b) {
      galic::CommitArgs C;
      memset(&C, 0, sizeof(C));
      C.n = b - a;
      C.K = K;
      C.trial_base = a;
      C.slot_pos = E.slot_pos;
      C.accept = c.t_accept;
      C.tidx = c.t_idx;
      C.ptype = c.P.type;
      return C;
    };
    // The reps of the group captured in one graph (the 16x16-tile form) and launched once:
    // nothing is waited for between the kernels of the reps.  The commits' bucket room comes from
    // a bound read back once per group (k_commit_bound), the rest of the host's work per rep
    // (the bucket layout, the launch order) is done on the device.
    int nrep_g = 0, max_n = 0;
    for(int r = r0; r < nrep && r * rep_size < s1; r++) {
      nrep_g++;
      max_n = std::max(max_n, std::min({ntrial, (r + 1) * rep_size, s1}) - std::max(r * rep_size, s0));
    }
    int m_order = 1;
    while(m_order < max_n) m_order <<= 1;
    if(t16 && c.use_graph && c.cuts_ready && max_n <= galic::kOrderMax &&
       sizeof(unsigned long long) * m_order <= lds_max) {
      if(c.cm_bound_len < nrep_g) {
        dfree(c.cm_bound);
        c.cm_bound_len = 0;
        CHK(dalloc(&c.cm_bound, (size_t)nrep_g));
        c.cm_bound_len = nrep_g;
      }
      hipLaunchKernelGGL(galic::k_fill_u32, dim3(nblocks(2 * (size_t)nrep_g, 256)), dim3(256), 0, c.stream,
                         (unsigned int *)c.cm_bound, 0u, 2 * (size_t)nrep_g);
      hipLaunchKernelGGL(galic::k_commit_bound, dim3(nblocks(ng, 256)), dim3(256), 0, c.stream, ng, K, s0, rep_size,
                         (const int *)c.t_inv, (const int *)c.log_n, c.cm_bound);
      HIPCHK(hipGetLastError());
      std::vector<unsigned long long> bound(nrep_g);
      HIPCHK(hipMemcpyAsync(bound.data(), c.cm_bound, sizeof(unsigned long long) * nrep_g, hipMemcpyDeviceToHost,
                            c.stream));
      HIPCHK(hipStreamSynchronize(c.stream));
      size_t room = 1;
      for(unsigned long long v : bound) room = std::max(room, (size_t)v);
      if(room <= 0x7fffffffu) {
        size_t cap_chunks = 0;
        CHK(ensure_commit_room(c, room, &cap_chunks));
        int ncu = 256;
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device);
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        int rc = GALIC_OK;
        hipLaunchKernelGGL(galic::k_fill_u32, dim3(1), dim3(64), 0, c.stream, c.cm_aux, 0u, (size_t)2);
        for(int r = r0; rc == GALIC_OK && r < nrep && r * rep_size < s1; r++) {
          const int ra = r * rep_size, rb = std::min(ntrial, ra + rep_size);
          const int a = std::max(ra, s0), b = std::min(rb, s1);
          if(!packed) {
            hipLaunchKernelGGL(galic::k_pack_merit, dim3(nblocks(nbw, 256)), dim3(256), 0, c.stream,
                               (const double *)c.stacks, (const double *)c.weights, c.merit_bw, c.merit_bwf, c.Nstack);
            packed = true;
          }
          galic::EvalArgs32 E = eval_args(a);
          rc = t16_rep(E, a, b, true);
          if(rc == GALIC_OK) rc = commit_logs_launch(c, commit_args(E, a, b), c.locA_dq, room, cap_chunks, ncu);
          if(rc == GALIC_OK && b == rb) {
            rc = commit_q(c);
            if(rc == GALIC_OK) rc = global_fit_launch(ctx);
            packed = false;
          }
        }
        const hipError_t ec = hipStreamEndCapture(c.stream, &graph);
        if(rc != GALIC_OK || ec != hipSuccess) {
          if(graph) (void)hipGraphDestroy(graph);
          return rc != GALIC_OK ? rc
                                : set_err(GALIC_ERR_HIP, "capturing the reps of a group failed: %s",
                                          hipGetErrorString(ec));
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if(ei != hipSuccess)
          return set_err(GALIC_ERR_HIP, "instantiating the reps' graph failed: %s", hipGetErrorString(ei));
        hipError_t el = hipEventRecord(c.ev0, c.stream);
        if(el == hipSuccess) el = hipGraphLaunch(exec, c.stream);
        if(el == hipSuccess) el = hipEventRecord(c.ev1, c.stream);
        if(el == hipSuccess) el = hipEventSynchronize(c.ev1);
        (void)hipGraphExecDestroy(exec);
        if(el != hipSuccess) return set_err(GALIC_ERR_HIP, "the reps' graph failed: %s", hipGetErrorString(el));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        eval_ms += ms;
        c.acc_eval_ms += ms;
        c.acc_graph_launches++;
        unsigned int terr = 0, cerr = 0;
        HIPCHK(hipMemcpyAsync(&terr, c.tile_aux + 1, sizeof(terr), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipMemcpyAsync(&cerr, c.cm_aux + 1, sizeof(cerr), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        if(terr) return set_err(GALIC_ERR_STATE, "the tile lists of a particle outgrew their room");
        if(cerr) return set_err(GALIC_ERR_STATE, "a commit outgrew its bound on bucket entries");
        return GALIC_OK;
      }
    }
    // evaluate the reps of the group in turn, committing each completed rep
    for(int r = r0; r < nrep && r * rep_size < s1; r++) {
      const int ra = r * rep_size, rb = std::min(ntrial, ra + rep_size);
      const int a = std::max(ra, s0), b = std::min(rb, s1);
      if(!packed) {
        hipLaunchKernelGGL(galic::k_pack_merit, dim3(nblocks(nbw, 256)), dim3(256), 0, c.stream,
                           (const double *)c.stacks, (const double *)c.weights, c.merit_bw, c.merit_bwf, c.Nstack);
        HIPCHK(hipGetLastError());
        packed = true;
      }
      galic::EvalArgs32 E = eval_args(a);
      HIPCHK(hipEventRecord(c.ev0, c.stream));
      if(t16) {
        CHK(t16_rep(E, a, b, false));
      } else if(tiles) {
        E.nblk = b - a;
        E.tlist = c.tile_list;
        E.tlist_cap = tile_cap;
        E.tile_err = c.tile_aux + 1;
        E.prof = c.merit_prof;
        E.tbl_lds = sparse_tbl() ? 1 : 0;
        const int grid = std::min(b - a, tile_grid);
        E.vinst = c.t_vinst;
        if(sparse) {
          // instantaneous terms before, the write-back of the choice after the merit kernel
          hipLaunchKernelGGL(galic::k_trial_inst, dim3(nblocks((size_t)(b - a) * K, 256)), dim3(256), 0, c.stream, c.M,
                             E, c.t_vinst);
          if(c.L == 9)
            hipLaunchKernelGGL((galic::k_eval_sparse32<9>), dim3(grid), dim3(1024), shmem, c.stream, c.M, E);
          else
            hipLaunchKernelGGL((galic::k_eval_sparse32<0>), dim3(grid), dim3(1024), shmem, c.stream, c.M, E);
          hipLaunchKernelGGL(galic::k_apply_best, dim3(nblocks(b - a, 256)), dim3(256), 0, c.stream, c.M, E);
        } else if(c.L == 9 && bwl)
          hipLaunchKernelGGL((galic::k_eval_tile32<9, true>), dim3(grid), dim3(1024), shmem, c.stream, c.M, E);
        else if(c.L == 9)
          hipLaunchKernelGGL((galic::k_eval_tile32<9, false>), dim3(grid), dim3(1024), shmem, c.stream, c.M, E);
        else if(bwl)
          hipLaunchKernelGGL((galic::k_eval_tile32<0, true>), dim3(grid), dim3(1024), shmem, c.stream, c.M, E);
        else
          hipLaunchKernelGGL((galic::k_eval_tile32<0, false>), dim3(grid), dim3(1024), shmem, c.stream, c.M, E);
        HIPCHK(hipGetLastError());
      } else if(gmem_now) {
        for(int b0 = 0; b0 < b - a; b0 += kGBlocks) {
          E.block_base = b0;
          if(c.L == 9)
            hipLaunchKernelGGL((galic::k_eval_log32<true, true, 9>), dim3(std::min(kGBlocks, b - a - b0)),
                               dim3(galic::merit_threads(9)), shmem, c.stream, c.M, E);
          else if(c.L == 8)
            hipLaunchKernelGGL((galic::k_eval_log32<true, true, 8>), dim3(std::min(kGBlocks, b - a - b0)),
                               dim3(galic::merit_threads(8)), shmem, c.stream, c.M, E);
          else
            hipLaunchKernelGGL((galic::k_eval_log32<true, true, 0>), dim3(std::min(kGBlocks, b - a - b0)), dim3(512),
                               shmem, c.stream, c.M, E);
          HIPCHK(hipGetLastError());
        }
      } else if(use_o) {
        // the common grid levels with their index arithmetic fixed at compile time
        if(c.L == 6)
          hipLaunchKernelGGL((galic::k_eval_log32<true, false, 6>), dim3(b - a), dim3(512), shmem, c.stream, c.M, E);
        else
          hipLaunchKernelGGL((galic::k_eval_log32<true, false, 0>), dim3(b - a), dim3(512), shmem, c.stream, c.M, E);
      } else {
        if(c.L == 7)
          hipLaunchKernelGGL((galic::k_eval_log32<false, false, 7>), dim3(b - a), dim3(galic::merit_threads(7)), shmem,
                             c.stream, c.M, E);
        else
          hipLaunchKernelGGL((galic::k_eval_log32<false, false, 0>), dim3(b - a), dim3(512), shmem, c.stream, c.M, E);
      }
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(c.ev1, c.stream));
      HIPCHK(hipEventSynchronize(c.ev1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
      eval_ms += ms;
      c.acc_eval_ms += ms;
      if(tiles || t16) {
        unsigned int terr = 0;
        HIPCHK(hipMemcpyAsync(&terr, c.tile_aux + 1, sizeof(terr), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        if(terr) return set_err(GALIC_ERR_STATE, "the tile lists of a particle outgrew their room");
      }
      {
        // the taken trials' orbits into the Loc deltas, the orbits they replace out (timed
        // with the merit evaluation)
        galic::CommitArgs C = commit_args(E, a, b);
        HIPCHK(hipEventRecord(c.ev0, c.stream));
        CHK(commit_logs(c, C, c.locA_dq));
        HIPCHK(hipEventRecord(c.ev1, c.stream));
        HIPCHK(hipEventSynchronize(c.ev1));
        float cms = 0;
        HIPCHK(hipEventElapsedTime(&cms, c.ev0, c.ev1));
        eval_ms += cms;
        c.acc_eval_ms += cms;
      }
      if(b == rb) {
        CHK(commit_q(c));
        CHK(global_fit(ctx));
        packed = false;
      }
    }
    return GALIC_OK;
  };
  // Group after group.  With two orbit sets, the next group's orbits integrate on the other set
  // while a group is judged, when the two groups' particles are distinct: the next group's trial
  // draws and orbits then depend on nothing the judging changes, so the results are the bits of
  // one group after the other.
  std::vector<char> in_group(overlap ? c.P.n : 0, 0);
  galic::OrbitArgs32 A_cur, A_next;
  int s0 = 0, s1 = 0;
  if(ntrial > 0) {
    s1 = plan(0);
    CHK(start(0, s1, A_cur));
  }
  while(s0 < ntrial) {
    CHK(finish(s0, s1, A_cur));
    cap = group_cap();
    int n0 = s1, n1 = s1;
    bool ahead = false;
    if(overlap && n0 < ntrial) {
      for(int i = s0; i < s1; i++) in_group[idx[i]] = 1;
      n1 = plan(n0);
      bool distinct = true;
      for(int i = n0; i < n1 && distinct; i++) distinct = !in_group[idx[i]];
      for(int i = s0; i < s1; i++) in_group[idx[i]] = 0;
      if(distinct) {
        swap_orbit_set(c);
        c.fp32_bpc = bpc_beside;
        const int rc = start(n0, n1, A_next);
        c.fp32_bpc = bpc_alone;
        swap_orbit_set(c);
        CHK(rc);
        ahead = true;
      }
    }
    CHK(judge(s0, s1));
    if(n0 < ntrial) {
      if(ahead) {
        swap_orbit_set(c);
        A_cur = A_next;
      } else {
        n1 = plan(n0);
        CHK(start(n0, n1, A_cur));
      }
    }
    s0 = n0;
    s1 = n1;
  }
  if(c.set_parity) swap_orbit_set(c);  // the first set active again
  if(ntrial == 0) {
    CHK(commit_q(c));
    CHK(global_fit(ctx));
  }
  c.last_eval_ms = eval_ms;
  if(ntrial > 0 && accept) HIPCHK(hipMemcpy(accept, c.t_accept, sizeof(int) * ntrial, hipMemcpyDeviceToHost));
  return GALIC_OK;
}

}  // namespace

int galic_fp32_merit_profile(galic_ctx *ctx, int on, unsigned long long *out16) {
  if(!ctx) return set_err(GALIC_ERR_ARG, "null context");
  galic::Ctx &c = ctx->c;
  HIPCHK(hipSetDevice(c.device));
  if(out16 && c.merit_prof) HIPCHK(hipMemcpy(out16, c.merit_prof, 16 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  if(on && !c.merit_prof) CHK(dalloc(&c.merit_prof, 16));
  if(on) HIPCHK(hipMemset(c.merit_prof, 0, 16 * sizeof(unsigned long long)));
  if(!on) dfree(c.merit_prof);
  return GALIC_OK;
}

int galic_fp32_set_merit_form(galic_ctx *ctx, int form) {
  if(!ctx || form < 0 || form > 4) return set_err(GALIC_ERR_ARG, "merit form must be 0, 1, 2, 3 or 4");
  ctx->c.merit_form = form;
  return GALIC_OK;
}

int galic_fp32_set_schedule(galic_ctx *ctx, int max_group_trials, int overlap, int graph) {
  if(!ctx || max_group_trials < 0 || overlap < -1 || overlap > 1 || graph < 0 || graph > 1)
    return set_err(GALIC_ERR_ARG, "max_group_trials must be >= 0, overlap -1, 0 or 1 and graph 0 or 1");
  ctx->c.max_group_trials = max_group_trials;
  ctx->c.overlap = overlap;
  ctx->c.use_graph = graph;
  return GALIC_OK;
}

int galic_fp32_set_merit_launch(galic_ctx *ctx, int wg_per_cu) {
  if(!ctx || wg_per_cu < 0 || wg_per_cu > 32) return set_err(GALIC_ERR_ARG, "wg_per_cu must be 0..32");
  ctx->c.merit_wg_per_cu = wg_per_cu;
  return GALIC_OK;
}

int galic_fp32_set_fresh_trials(galic_ctx *ctx, int nfresh) {
  if(!ctx || nfresh < 0 || nfresh > 64) return set_err(GALIC_ERR_ARG, "nfresh must be 0..64");
  ctx->c.fp32_nfresh = nfresh;
  return GALIC_OK;
}

int galic_fp32_rep(galic_ctx *ctx, int ntrial, const int *idx, int K, const double *vtry, double sigma_frac,
                   unsigned long long seed, int *accept) {
  return fp32_run(ctx, ntrial, idx, K, std::max(ntrial, 1), vtry, sigma_frac, seed, accept);
}

int galic_fp32_iterate(galic_ctx *ctx, int ntrial, const int *idx, int K, int rep_size, const double *vtry,
                       double sigma_frac, unsigned long long seed, int *accept) {
  return fp32_run(ctx, ntrial, idx, K, rep_size, vtry, sigma_frac, seed, accept);
}

int galic_fp32_autotune_schedule(galic_ctx *ctx, int ntrial, const int *idx, int K, int rep_size,
                                 double sigma_frac, unsigned long long seed, int *accept, int *orbit_wg_per_cu,
                                 int *merit_wg_per_cu) {
  if(!ctx || !idx || ntrial < 1 || rep_size < 1) return set_err(GALIC_ERR_ARG, "bad schedule-tuning arguments");
  galic::Ctx &c = ctx->c;
  if(!c.fp32_tuned) return set_err(GALIC_ERR_STATE, "galic_fp32_autotune must run first");
  // (orbit kernel workgroups per CU beside the judging, merit kernel workgroups per CU (0: all))
  const int base = c.fp32_bpc;
  const int cand[3][2] = {{std::max(1, base / 2), 0}, {base, 0}, {std::max(1, base / 2), 1}};
  const int ncand = 3;
  // two runs per candidate, interleaved, the better of the two counted (the first run of all
  // also pays the call's first allocations); the last run takes the reps left over
  const int nrep = (ntrial + rep_size - 1) / rep_size;
  const int per = std::max(1, nrep / (2 * ncand));
  double rate[3] = {0, 0, 0};
  for(int i = 0, r = 0; r < nrep; i++) {
    const int r1 = i < 2 * ncand - 1 ? std::min(nrep, r + per) : nrep;
    const int a = r * rep_size, b = std::min(ntrial, r1 * rep_size);
    const int ci = i % ncand;
    c.fp32_bpc_overlap = cand[ci][0];
    c.merit_wg_per_cu = cand[ci][1];
    const auto t0 = std::chrono::steady_clock::now();
    CHK(fp32_run(ctx, b - a, idx + a, K, rep_size, nullptr, sigma_frac, seed, accept ? accept + a : nullptr));
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double rt = (double)(b - a) / std::max(sec, 1e-9);  // trials per second
    if(getenv("GALIC_VERBOSE"))
      fprintf(stderr, "[galic] schedule tuning: orbit %d WG/CU beside the judging, merit %d WG/CU (0: all): "
                      "%d reps, %.4g trials/s\n", cand[ci][0], cand[ci][1], r1 - r, rt);
    rate[ci] = std::max(rate[ci], rt);
    r = r1;
  }
  int bi = 1;  // (the autotuned launch when no candidate was measured)
  for(int i = 0; i < ncand; i++)
    if(rate[i] > rate[bi]) bi = i;
  const int bo = cand[bi][0], bm = cand[bi][1];
  c.fp32_bpc_overlap = bo;
  c.merit_wg_per_cu = bm;
  if(orbit_wg_per_cu) *orbit_wg_per_cu = bo;
  if(merit_wg_per_cu) *merit_wg_per_cu = bm;
  return GALIC_OK;
}

int galic_fp32_trial_merits(galic_ctx *ctx, int ntrial, int k, double *dFk) {
  if(!ctx || !dFk) return set_err(GALIC_ERR_ARG, "null argument");
  galic::Ctx &c = ctx->c;
  if(ntrial != c.last_ntrial || k != c.last_k || ntrial <= 0)
    return set_err(GALIC_ERR_STATE, "the last FP32 call had %d trials of K=%d", c.last_ntrial, c.last_k);
  HIPCHK(hipSetDevice(c.device));
  HIPCHK(hipMemcpy(dFk, c.t_dFk, sizeof(double) * (size_t)ntrial * k, hipMemcpyDeviceToHost));
  return GALIC_OK;
}

int galic_timings(galic_ctx *ctx, double *out11) {
  if(!ctx || !out11) return set_err(GALIC_ERR_ARG, "null argument");
  galic::Ctx &c = ctx->c;
  out11[0] = c.last_orbit_ms;
  out11[1] = c.last_eval_ms;
  out11[2] = c.last_fit_ms;
  out11[3] = c.acc_orbit_ms;
  out11[4] = (double)c.acc_orbit_launches;
  out11[5] = (double)c.acc_orbits;
  out11[6] = (double)c.acc_steps;
  out11[7] = c.acc_eval_ms;
  out11[8] = c.acc_fit_ms;
  out11[9] = (double)c.acc_log_relaunches;
  out11[10] = (double)c.acc_graph_launches;
  return GALIC_OK;
}

int galic_reset_timings(galic_ctx *ctx) {
  if(!ctx) return set_err(GALIC_ERR_ARG, "null argument");
  galic::Ctx &c = ctx->c;
  c.acc_orbit_ms = 0;
  c.acc_orbit_launches = 0;
  c.acc_orbits = 0;
  c.acc_steps = 0;
  c.acc_eval_ms = 0;
  c.acc_fit_ms = 0;
  c.acc_log_relaunches = 0;
  c.acc_graph_launches = 0;
  return GALIC_OK;
}

int galic_bind(galic_ctx *ctx) {
  g_bound = ctx;
  return GALIC_OK;
}

}  // extern "C"
