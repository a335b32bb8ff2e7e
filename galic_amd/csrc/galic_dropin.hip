// galic_dropin.hip -- the reference's own entry points (proto.h:100-145) on
// top of the batched device paths.  Each call works on the context made current
// with galic_bind(), the way the reference's functions work on its globals.
// They are correct but launch-bound: one orbit or one point per call.  A C host
// that can batch should call galic_orbit_responses / galic_field_at instead.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/galic_hip.h"
#include "galic_ctx.h"
#include "galic_kernels.h"

// defined in galic_api.hip
struct galic_ctx;
namespace galic {
galic_ctx *bound_ctx();
Ctx &ctx_state(galic_ctx *ctx);
}  // namespace galic

namespace galic {

// forcegrid_get_acceleration and forcegrid_get_potential at n points (grid.c:59-197)
__global__ void k_field_at(DevModel M, int n, const double *pos, double *acc, double *pot) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
  FgIndex64 g = fg_index64(M, x, y, z);
  double ax, ay, az;
  accel64(M, x, y, z, g, ax, ay, az);
  if(acc) {
    acc[3 * i] = ax;
    acc[3 * i + 1] = ay;
    acc[3 * i + 2] = az;
  }
  if(pot) pot[i] = potential64(M, x, y, z);
}

// smooth_stack (grid.c:1247-1265) in one workgroup
__global__ void k_smooth_one(double *data, int L) {
  for(int l = L - 1; l >= 0; l--) {
    int n = 1 << l;
    for(int q = threadIdx.x; q < n * n; q += blockDim.x) {
      int i = q / n, j = q % n;
      data[GALIC_STACKOFFSET(l, i, j)] =
          data[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j)] + data[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j)] +
          data[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j + 1)] + data[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j + 1)];
    }
    __syncthreads();
  }
}

}  // namespace galic

static void dropin_fail(const char *what) {
  // The reference terminates on errors inside these functions (allvars.h:129,
  // orbit_response.c:449-456); the drop-ins do the same.
  fprintf(stderr, "galic drop-in %s failed: %s\n", what, galic_last_error());
  exit(1);
}

extern "C" {

int galic_field_at(galic_ctx *ctx, int n, const double *pos, double *acc, double *pot) {
  if(!ctx || n < 0 || (n > 0 && !pos)) return GALIC_ERR_ARG;
  if(n == 0) return GALIC_OK;
  galic::Ctx &c = galic::ctx_state(ctx);
  if(hipSetDevice(c.device) != hipSuccess) return GALIC_ERR_HIP;
  double *dpos = nullptr, *dacc = nullptr, *dpot = nullptr;
  int rc = GALIC_OK;
  if(hipMalloc(&dpos, sizeof(double) * 3 * n) != hipSuccess || hipMalloc(&dacc, sizeof(double) * 3 * n) != hipSuccess ||
     hipMalloc(&dpot, sizeof(double) * n) != hipSuccess)
    rc = GALIC_ERR_NOMEM;
  if(rc == GALIC_OK && hipMemcpy(dpos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice) != hipSuccess)
    rc = GALIC_ERR_HIP;
  if(rc == GALIC_OK) {
    hipLaunchKernelGGL(galic::k_field_at, dim3((n + 255) / 256), dim3(256), 0, c.stream, c.M, n, dpos, dacc, dpot);
    if(hipGetLastError() != hipSuccess || hipStreamSynchronize(c.stream) != hipSuccess) rc = GALIC_ERR_HIP;
  }
  if(rc == GALIC_OK && acc && hipMemcpy(acc, dacc, sizeof(double) * 3 * n, hipMemcpyDeviceToHost) != hipSuccess)
    rc = GALIC_ERR_HIP;
  if(rc == GALIC_OK && pot && hipMemcpy(pot, dpot, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    rc = GALIC_ERR_HIP;
  (void)hipFree(dpos);
  (void)hipFree(dacc);
  (void)hipFree(dpot);
  return rc;
}

double produce_orbit_response_field_mod(double *pos, double *vel, int id, double *mfield, double *egyfield_r,
                                        double *egyfield_t, double *egyfield_q, double *egyfield_p, double mass,
                                        double timespan, int *orbitstaken, int type) {
  (void)id;
  galic_ctx *ctx = galic::bound_ctx();
  if(!ctx) dropin_fail("produce_orbit_response_field_mod (no context bound)");
  galic::Ctx &c = galic::ctx_state(ctx);
  std::vector<double> f((size_t)galic::kFields * c.Ngrid);
  double t = 0;
  int norb = 0;
  if(galic_orbit_responses(ctx, 1, pos, vel, &mass, &timespan, &type, f.data(), &t, &norb, nullptr, nullptr,
                           nullptr) != GALIC_OK)
    dropin_fail("produce_orbit_response_field_mod");
  const size_t ng = c.Ngrid;
  memcpy(mfield, f.data(), sizeof(double) * ng);
  memcpy(egyfield_r, f.data() + ng, sizeof(double) * ng);
  memcpy(egyfield_t, f.data() + 2 * ng, sizeof(double) * ng);
  memcpy(egyfield_q, f.data() + 3 * ng, sizeof(double) * ng);
  memcpy(egyfield_p, f.data() + 4 * ng, sizeof(double) * ng);
  *orbitstaken = norb;
  return t;
}

double produce_orbit_response_field(double *pos, double *vel, int id, double *mfield, double mass, double timespan,
                                    int *orbitstaken) {
  // The plain variant bins only the mass field; its orbit and mass field are
  // those of the _mod variant (orbit_response.c:33-151 vs 155-491).
  galic_ctx *ctx = galic::bound_ctx();
  if(!ctx) dropin_fail("produce_orbit_response_field (no context bound)");
  galic::Ctx &c = galic::ctx_state(ctx);
  std::vector<double> e((size_t)4 * c.Ngrid);
  const size_t ng = c.Ngrid;
  return produce_orbit_response_field_mod(pos, vel, id, mfield, e.data(), e.data() + ng, e.data() + 2 * ng,
                                          e.data() + 3 * ng, mass, timespan, orbitstaken, 1);
}

double forcegrid_get_potential(double *pos) {
  galic_ctx *ctx = galic::bound_ctx();
  double pot = 0;
  if(!ctx || galic_field_at(ctx, 1, pos, nullptr, &pot) != GALIC_OK) dropin_fail("forcegrid_get_potential");
  return pot;
}

double forcegrid_get_escape_speed(double *pos) {
  double phi = forcegrid_get_potential(pos);
  return sqrt(-2.0 * phi);  // grid.c:115-120
}

void forcegrid_get_acceleration(double *pos, double *acc) {
  galic_ctx *ctx = galic::bound_ctx();
  if(!ctx || galic_field_at(ctx, 1, pos, acc, nullptr) != GALIC_OK) dropin_fail("forcegrid_get_acceleration");
}

void smooth_stack(double *data, int maxlevel) {
  if(maxlevel < 0 || maxlevel > 12) dropin_fail("smooth_stack (maxlevel)");
  size_t n = GALIC_NSTACK(maxlevel);
  double *d = nullptr;
  if(hipMalloc(&d, sizeof(double) * n) != hipSuccess) dropin_fail("smooth_stack (alloc)");
  if(hipMemcpy(d, data, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) dropin_fail("smooth_stack");
  hipLaunchKernelGGL(galic::k_smooth_one, dim3(1), dim3(256), 0, 0, d, maxlevel);
  if(hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) dropin_fail("smooth_stack");
  if(hipMemcpy(data, d, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) dropin_fail("smooth_stack");
  (void)hipFree(d);
}

}  // extern "C"
