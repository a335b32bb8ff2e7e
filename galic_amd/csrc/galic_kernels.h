// galic_kernels.h -- kernel argument blocks and launch declarations shared by
// the FP64 core, the FP32 batched path and the host API.
#pragma once
#include <hip/hip_runtime.h>

#include "galic_ctx.h"

#define GALIC_OK 0
#define GALIC_ERR_ARG (-1)
#define GALIC_ERR_HIP (-2)
#define GALIC_ERR_STEPCAP (-3)
#define GALIC_ERR_STATE (-4)
#define GALIC_ERR_NOMEM (-5)

namespace galic {

// Records the message returned by galic_last_error(); returns code.
int set_error(int code, const char *msg);

struct OrbitArgs {
  int n;               // number of orbits in this launch
  unsigned int *counter;
  const double *ox, *oy, *oz;     // start position of each orbit
  const double *ovx, *ovy, *ovz;  // start velocity of each orbit
  const double *omass, *otint;    // mass and integration time of each orbit
  const int *otype;
  double *slots;       // private response per orbit (trial evaluation), or null
  size_t slot_stride;  // doubles per slot (5*Ngrid)
  int slot_base;       // orbit index of slot 0
  double *loc;         // Loc grids [type][5][Ngrid] for direct atomics, or null
  double *tall, *e0, *e1;
  int *norbit, *status;
  long long *steps;
  int4 *bbox;          // touched cells: R from .x to .y, z from .z to .w (inclusive)
  unsigned long long *stats;  // [0] += steps, [1] = failed orbit + 1, [2] = its status
};

struct EvalArgs {
  Particles P;
  const int *tidx;
  const double *tvx, *tvy, *tvz;
  const double *stacks, *weights, *fit;
  const int *cutB_of;  // [type][Ngrid] -> cut-B node of each fine cell
  const double *slots;
  size_t slot_stride;
  int slot_base;
  const int4 *bbox;
  const int *norbit;
  double *locA_d, *locI_d;
  int *accept;
  double *dF;
  double MType[GALIC_NTYPES];
  double fac[GALIC_NTYPES][4];  // fac_value rs, ts, qs, ps
  double lowdisp;
  int block_base;      // trial of the launch's first workgroup
  double *cs_global;   // coarse-level scratch in HBM, cs_stride doubles per workgroup (grids too large for LDS), or null
  size_t cs_stride;
};

struct FitArgs {
  int L;
  int present[GALIC_NTYPES];
  int tovs[GALIC_NTYPES];
  double *stacks;
  const double *weights;
  const double *locA, *locI, *locM;
  double *fit;
  double lowdisp;
};

__global__ void k_orbit64(DevModel M, OrbitArgs A);

// Per-lane totals of an orbit launch folded into three counters (no per-orbit download).
__device__ __forceinline__ void orbit_stats_add(unsigned long long *stats, long long steps, int status, int o) {
  atomicAdd(&stats[0], (unsigned long long)steps);
  if(status != 0 && atomicCAS(&stats[1], 0ull, (unsigned long long)o + 1) == 0ull)
    atomicExch(&stats[2], (unsigned long long)(long long)status);
}
__global__ void k_inst_init(DevModel M, Particles P, double *locI, double *locM, int ngrid);
__global__ void k_eval_apply64(DevModel M, EvalArgs A);
__global__ void k_commit(double *loc, double *delta, size_t n);
__global__ void k_fit_stacks(FitArgs A);
__global__ void k_fit_sums(FitArgs A);
__global__ void k_fit_tiles(FitArgs A);
__global__ void k_fit_top(FitArgs A);
__global__ void k_fit_partial(FitArgs A, double *part, int nblk);
__global__ void k_fit_final(FitArgs A, const double *part, int nblk);
constexpr int kFitBlocks = 64;  // workgroups per type of the fit's cut sums
struct OrbitIn {
  int *pidx;
  double *ox, *oy, *oz, *ovx, *ovy, *ovz, *omass, *otint;
  int *otype;
};
__global__ void k_prepare_trials(int ntrial, const int *tidx, const double *tvx, const double *tvy,
                                 const double *tvz, Particles P, OrbitIn O);
__global__ void k_prepare_all(int n, int base, Particles P, OrbitIn O);
__global__ void k_store_orbits(int n, int base, const int *norbit, int *orbits);

}  // namespace galic
