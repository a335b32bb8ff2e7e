// galic_noalloc.hip -- galic_field_at_static: the force-field lookups of galic_field_at
// (forcegrid_get_acceleration / forcegrid_get_potential, grid.c:59-197) with their device
// staging in a static array of this library instead of memory the call allocates, for
// callers that cannot allocate device memory (code under HIP graph capture, tight loops).
//
// The array holds GALIC_FIELD_STATIC_MAX points and is shared by every call on one GPU:
// calls that overlap on one GPU race on it.  Each GPU keeps its own copy until the process
// ends; its device address is looked up once per GPU, at the first call there.
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <mutex>

#include "../../include/galic_hip.h"
#include "galic_ctx.h"

namespace galic {
Ctx &ctx_state(galic_ctx *ctx);
int set_error(int code, const char *msg);
int field_lookup(Ctx &c, int n, const double *pos, double *acc, double *pot, double *dpos, double *dacc,
                 double *dpot);
}  // namespace galic

// staging: positions [3 * max], accelerations [3 * max], potentials [max]
static __device__ __align__(256) double scratch[7 * GALIC_FIELD_STATIC_MAX];

namespace {
constexpr int kMaxDevices = 64;
void *scratch_dev[kMaxDevices];  // device address of `scratch` on each GPU, once looked up
std::mutex scratch_mu;

// The device address of `scratch` on the current GPU (dev), looked up at the first call.
int scratch_address(int dev, double **out) {
  std::lock_guard<std::mutex> lock(scratch_mu);
  if(!scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if(e != hipSuccess || !p) {
      char msg[160];
      snprintf(msg, sizeof msg, "galic_field_at_static: no device address of the static staging on GPU %d (%s)", dev,
               hipGetErrorString(e));
      return galic::set_error(GALIC_ERR_HIP, msg);
    }
    scratch_dev[dev] = p;
  }
  *out = (double *)scratch_dev[dev];
  return GALIC_OK;
}
}  // namespace

extern "C" int galic_field_at_static(galic_ctx *ctx, int n, const double *pos, double *acc, double *pot) {
  if(!ctx || n < 0 || (n > 0 && !pos)) return galic::set_error(GALIC_ERR_ARG, "galic_field_at_static: bad arguments");
  if(n > GALIC_FIELD_STATIC_MAX) {
    char msg[160];
    snprintf(msg, sizeof msg, "galic_field_at_static: %d points exceed the static staging of %d points", n,
             GALIC_FIELD_STATIC_MAX);
    return galic::set_error(GALIC_ERR_ARG, msg);
  }
  if(n == 0) return GALIC_OK;
  galic::Ctx &c = galic::ctx_state(ctx);
  if(c.device < 0 || c.device >= kMaxDevices) return galic::set_error(GALIC_ERR_ARG, "galic_field_at_static: device");
  if(hipSetDevice(c.device) != hipSuccess) return galic::set_error(GALIC_ERR_HIP, "galic_field_at_static: hipSetDevice");
  double *base = nullptr;
  const int rc = scratch_address(c.device, &base);
  if(rc != GALIC_OK) return rc;
  double *dpos = base, *dacc = base + 3 * (size_t)GALIC_FIELD_STATIC_MAX, *dpot = base + 6 * (size_t)GALIC_FIELD_STATIC_MAX;
  const int r = galic::field_lookup(c, n, pos, acc, pot, dpos, dacc, dpot);
  return r == GALIC_OK ? r : galic::set_error(r, "galic_field_at_static: HIP error in the lookup");
}
