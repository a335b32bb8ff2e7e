// galic_dropin.hip -- the reference's own entry points (proto.h:100-145) on
// top of the batched device paths.  Each call works on the context made current
// with galic_bind(), the way the reference's functions work on its globals.
// They are correct but launch-bound: one orbit or one point per call.  A C host
// that can batch should call galic_orbit_responses / galic_field_at instead.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/galic_hip.h"
#include "galic_ctx.h"
#include "galic_kernels.h"
#include "galic_handle.h"

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

// n lookups through device staging (positions dpos, accelerations dacc, potentials dpot):
// copy in, one launch, copy out, on the context's stream.
int field_lookup(Ctx &c, int n, const double *pos, double *acc, double *pot, double *dpos, double *dacc,
                 double *dpot) {
  if(hipMemcpyAsync(dpos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c.stream) != hipSuccess)
    return GALIC_ERR_HIP;
  hipLaunchKernelGGL(k_field_at, dim3((n + 255) / 256), dim3(256), 0, c.stream, c.M, n, (const double *)dpos, dacc,
                     dpot);
  if(hipGetLastError() != hipSuccess) return GALIC_ERR_HIP;
  if(acc && hipMemcpyAsync(acc, dacc, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c.stream) != hipSuccess)
    return GALIC_ERR_HIP;
  if(pot && hipMemcpyAsync(pot, dpot, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream) != hipSuccess)
    return GALIC_ERR_HIP;
  return hipStreamSynchronize(c.stream) == hipSuccess ? GALIC_OK : GALIC_ERR_HIP;
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
  // the context's staging, grown to the largest call: a lookup in a loop allocates once
  if((size_t)n > c.lookup_len) {
    if(c.lookup_buf) (void)hipFree(c.lookup_buf);
    c.lookup_buf = nullptr;
    c.lookup_len = 0;
    const size_t len = std::max<size_t>((size_t)n, 1024);
    if(hipMalloc(&c.lookup_buf, sizeof(double) * 7 * len) != hipSuccess) return GALIC_ERR_NOMEM;
    c.lookup_len = len;
  }
  double *dpos = c.lookup_buf, *dacc = dpos + 3 * c.lookup_len, *dpot = dacc + 3 * c.lookup_len;
  return galic::field_lookup(c, n, pos, acc, pot, dpos, dacc, dpot);
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

// Log-index cell finders: the force grid's nodes (grid.c:15-55) and the
// cell-centred density and energy grids (grid.c:494-534, 896-938), on the host
// in the reference's operation order.
static void log_cell(const double *pos, double rmin, double fac, int nbin, double shift, int *iR, int *iz, double *fR,
                     double *fz) {
  double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
  double z = fabs(pos[2]);
  double br = (log(r / rmin + 1.0) / log(fac)) - shift;
  double bz = (log(z / rmin + 1.0) / log(fac)) - shift;
  if(br < 0) br = 0;
  if(bz < 0) bz = 0;
  int binr = (int)br, binz = (int)bz;
  *fR = br - binr;
  *fz = bz - binz;
  if(binr >= nbin - 1) {
    binr = nbin - 2;
    *fR = 1;
  }
  if(binz >= nbin - 1) {
    binz = nbin - 2;
    *fz = 1;
  }
  *iR = binr;
  *iz = binz;
}

static const galic_params &bound_params(const char *what) {
  galic_ctx *ctx = galic::bound_ctx();
  if(!ctx) dropin_fail(what);
  return galic::ctx_state(ctx).p;
}

void forcegrid_get_cell(double *pos, int *iR, int *iz, double *fR, double *fz) {
  const galic_params &p = bound_params("forcegrid_get_cell (no context bound)");
  log_cell(pos, p.FG_Rmin, p.FG_Fac, p.FG_Nbin, 0.0, iR, iz, fR, fz);
}

void densitygrid_get_cell(double *pos, int *iR, int *iz, double *fR, double *fz) {
  const galic_params &p = bound_params("densitygrid_get_cell (no context bound)");
  log_cell(pos, p.DG_Rmin, p.DG_Fac, 1 << p.DG_MaxLevel, 0.5, iR, iz, fR, fz);
}

void energygrid_get_cell(double *pos, int *iR, int *iz, double *fR, double *fz) {
  const galic_params &p = bound_params("energygrid_get_cell (no context bound)");
  log_cell(pos, p.EG_Rmin, p.EG_Fac, 1 << p.EG_MaxLevel, 0.5, iR, iz, fR, fz);
}

// orbit_response.c:16-26: DG_CellSize of the bound context's density grid
double get_timestep(double *pos, double *vel, double *acc, int icell) {
  (void)pos;
  galic_ctx *ctx = galic::bound_ctx();
  if(!ctx) dropin_fail("get_timestep (no context bound)");
  const galic_params &p = galic::ctx_state(ctx).p;
  if(icell < 0 || icell >= (int)ctx->cellsize.size()) dropin_fail("get_timestep (cell index)");
  double v = sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
  double aa = sqrt(acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2]);
  double torbit = p.V200 / aa;
  double tcross = ctx->cellsize[icell] / v;
  double a = p.TimeStepFactorOrbit * torbit, b = p.TimeStepFactorCellCross * tcross;
  return a < b ? a : b;
}

// calc_stack_difference (grid.c:1086-1112) and calc_stack_difference_mod
// (grid.c:1115-1151): the recursion stops at maxlevel or where ref1 < thresh
// and sums the term of `flag` there.
static double stack_term(const double *d1, const double *d2, int o, const double *ref1, const double *ref2,
                         const double *dist, int flag, int mod, const galic_params &p) {
  const double r20 = p.R200 / 20.0;
  const double dcap = dist[o] < r20 ? dist[o] : r20;  // min() of grid.c
  if(flag == 1 || (!mod && flag)) {
    if(ref1[o] > 0 && ref2[o] > 0) {
      const double q = d2[o] / ref2[o];
      return fabs(d1[o] / ref1[o] - d2[o] / ref2[o]) / (q > p.LowerDispLimit ? q : p.LowerDispLimit);  // dmax()
    }
    return 0;
  }
  if(mod && flag == 2) {
    if(ref2[o] > 0) return fabs(d1[o] - d2[o]) * (ref1[o] / ref2[o]) / dcap;
    return 0;
  }
  return fabs(d1[o] - d2[o]) / dcap;
}

static double stack_difference(const double *d1, const double *d2, int l, int i, int j, int maxlevel,
                               const double *ref1, const double *ref2, double thresh, const double *dist, int flag,
                               int mod, const galic_params &p) {
  const int o = GALIC_STACKOFFSET(l, i, j);
  if(l >= maxlevel || ref1[o] < thresh) return stack_term(d1, d2, o, ref1, ref2, dist, flag, mod, p);
  double sum = 0;
  for(int ii = 2 * i; ii <= 2 * i + 1; ii++)
    for(int jj = 2 * j; jj <= 2 * j + 1; jj++)
      sum += stack_difference(d1, d2, l + 1, ii, jj, maxlevel, ref1, ref2, thresh, dist, flag, mod, p);
  return sum;
}

double calc_stack_difference(double *d1, double *d2, int l, int i, int j, int maxlevel, double *ref1, double *ref2,
                             double thresh, double *dist, int flag) {
  const galic_params &p = bound_params("calc_stack_difference (no context bound)");
  return stack_difference(d1, d2, l, i, j, maxlevel, ref1, ref2, thresh, dist, flag, 0, p);
}

double calc_stack_difference_mod(double *d1, double *d2, int l, int i, int j, int maxlevel, double *ref1, double *ref2,
                                 double thresh, double *dist, int flag) {
  const galic_params &p = bound_params("calc_stack_difference_mod (no context bound)");
  return stack_difference(d1, d2, l, i, j, maxlevel, ref1, ref2, thresh, dist, flag, 1, p);
}

void smooth_stack(double *data, int maxlevel) {
  if(maxlevel < 0 || maxlevel > 12) dropin_fail("smooth_stack (maxlevel)");
  // a device buffer per GPU, grown to the largest stack and kept: repeated calls do not allocate
  static double *buf[64];
  static size_t len[64];
  int dev = 0;
  if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dropin_fail("smooth_stack (device)");
  const size_t n = GALIC_NSTACK(maxlevel);
  if(n > len[dev]) {
    if(buf[dev]) (void)hipFree(buf[dev]);
    buf[dev] = nullptr;
    len[dev] = 0;
    if(hipMalloc(&buf[dev], sizeof(double) * n) != hipSuccess) dropin_fail("smooth_stack (alloc)");
    len[dev] = n;
  }
  double *d = buf[dev];
  if(hipMemcpy(d, data, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) dropin_fail("smooth_stack");
  hipLaunchKernelGGL(galic::k_smooth_one, dim3(1), dim3(256), 0, 0, d, maxlevel);
  if(hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) dropin_fail("smooth_stack");
  if(hipMemcpy(data, d, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) dropin_fail("smooth_stack");
}

}  // extern "C"
