// galic_io.hip -- the reference's output files, written from the device state,
// so a run of this library can be compared file by file with a reference run:
//   fit_<type>.txt             one line per iteration (log_message, main.c:1221-1288)
//   densfield_<type>_<n>.dat   target and response maps (output_density_field, io.c:36-457)
//   snapshot (Gadget format 1) particle positions, velocities, IDs and masses
//                              (output_particles/write_file, io.c:460-1240, one file)
// Host code only: the grids and particles are copied back and formatted here.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "galic_handle.h"
#include "galic_kernels.h"

namespace {

int io_err(int code, const char *msg) { return galic::set_error(code, msg); }

#define IO_HIP(expr)                                        \
  do {                                                      \
    if((expr) != hipSuccess) return io_err(GALIC_ERR_HIP, #expr); \
  } while(0)

// eval_smoothed_stack (grid.c:1232-1244): the value of the deepest node on the
// path from fine cell (i, j) up whose parent's reference exceeds thresh, per fine cell.
double eval_smoothed(const double *din, int l, int i, int j, int maxlevel, const double *ref, double thresh) {
  while(l > 0 && !(ref[GALIC_STACKOFFSET(l - 1, i / 2, j / 2)] > thresh)) {
    l--;
    i /= 2;
    j /= 2;
  }
  const double n = (double)(1 << (maxlevel - l));
  return din[GALIC_STACKOFFSET(l, i, j)] / (n * n);
}

// calc_smoothed_stack (grid.c:1217-1229)
void smoothed_map(const double *din, double *dout, int maxlevel, const double *ref, double thresh) {
  const int n = 1 << maxlevel;
  for(int i = 0; i < n; i++)
    for(int j = 0; j < n; j++) dout[i * n + j] = eval_smoothed(din, maxlevel, i, j, maxlevel, ref, thresh);
}

struct File {
  FILE *f = nullptr;
  bool ok = true;
  ~File() {
    if(f) fclose(f);
  }
  template <typename T>
  void put(const T *p, size_t n) {
    if(ok && fwrite(p, sizeof(T), n, f) != n) ok = false;
  }
  void put_int(int v) { put(&v, 1); }
};

int stack_of(galic_ctx *ctx, int type, int which, std::vector<double> &out) {
  out.resize(ctx->c.Nstack);
  return galic_get_stack(ctx, type, which, out.data());
}

}  // namespace

extern "C" {

int galic_write_fit_line(galic_ctx *ctx, int type, const char *path, double opt_frac, long long tries,
                         long long changes) {
  if(!ctx || type < 1 || type > 3 || !path) return io_err(GALIC_ERR_ARG, "bad arguments");
  galic::Ctx &c = ctx->c;
  double fit[8];
  int rc = galic_get_fit(ctx, type, fit);
  if(rc != GALIC_OK) return rc;
  // vavg: rms speed of the type's particles (main.c:1236-1256)
  const int n = c.P.n;
  std::vector<double> vx(n), vy(n), vz(n);
  std::vector<int> t(n);
  IO_HIP(hipSetDevice(c.device));
  IO_HIP(hipStreamSynchronize(c.stream));
  if(n > 0) {
    IO_HIP(hipMemcpy(vx.data(), c.P.vx, n * sizeof(double), hipMemcpyDeviceToHost));
    IO_HIP(hipMemcpy(vy.data(), c.P.vy, n * sizeof(double), hipMemcpyDeviceToHost));
    IO_HIP(hipMemcpy(vz.data(), c.P.vz, n * sizeof(double), hipMemcpyDeviceToHost));
    IO_HIP(hipMemcpy(t.data(), c.P.type, n * sizeof(int), hipMemcpyDeviceToHost));
  }
  double vsum = 0;
  long long count = 0;
  for(int i = 0; i < n; i++)
    if(t[i] == type) {
      vsum += vx[i] * vx[i];
      vsum += vy[i] * vy[i];
      vsum += vz[i] * vz[i];
      count++;
    }
  if(count == 0) return GALIC_OK;  // the reference writes no line for an empty type
  const double vavg = sqrt(vsum / count);
  const double S = fit[0], sr = fit[1], st = fit[2], sp = fit[3], sq = fit[4], relc = fit[6];
  FILE *f = fopen(path, "a");
  if(!f) return io_err(GALIC_ERR_ARG, "cannot open the fit file for appending");
  const int w = fprintf(f, "%10.7g  %10.7g  %10.7g     %10.7g %10.7g %10.7g %10.7g     %10.7g    %10.7g\n", opt_frac, S,
                        vavg, sr, st, sp, sq, S + (sr + st + sp + sq) / relc,
                        ((double)changes) / (tries + 1.0e-60));
  const bool ok = w > 0 && fclose(f) == 0;
  return ok ? GALIC_OK : io_err(GALIC_ERR_ARG, "writing the fit file failed");
}

int galic_write_densfield(galic_ctx *ctx, int type, const char *path) {
  if(!ctx || type < 1 || type > 3 || !path) return io_err(GALIC_ERR_ARG, "bad arguments");
  galic::Ctx &c = ctx->c;
  const galic_params &p = c.p;
  if(!(p.NType[type] > 0)) return io_err(GALIC_ERR_ARG, "the type has no particles");
  if(!c.targets_set[type]) return io_err(GALIC_ERR_STATE, "the type's targets are not set");
  const int L = c.L, nb = 1 << L, ng = nb * nb;
  std::vector<double> dmt, dmr, emt, emr, et[4], er[4];
  int rc;
  if((rc = stack_of(ctx, type, 10, dmt)) || (rc = stack_of(ctx, type, 0, dmr)) || (rc = stack_of(ctx, type, 11, emt)) ||
     (rc = stack_of(ctx, type, 1, emr)))
    return rc;
  for(int k = 0; k < 4; k++)
    if((rc = stack_of(ctx, type, 12 + k, et[k])) || (rc = stack_of(ctx, type, 2 + k, er[k]))) return rc;
  const double thr_d = p.MinParticlesPerBinForDensityMeasurement * p.MType[type] / p.NType[type];
  const double thr_s = p.MinParticlesPerBinForDispersionMeasurement * p.MType[type] / p.NType[type];
  File F;
  F.f = fopen(path, "wb");
  if(!F.f) return io_err(GALIC_ERR_ARG, "cannot create the densfield file");
  std::vector<double> dout(ng);
  F.put_int(L);
  for(int lev = L; lev >= 0; lev--) {  // the target stack, finest level first
    F.put_int(1 << lev);
    F.put(&dmt[GALIC_STACKOFFSET(lev, 0, 0)], (size_t)1 << (2 * lev));
  }
  smoothed_map(dmt.data(), dout.data(), L, dmt.data(), thr_d);  // target, variable resolution
  F.put_int(nb);
  F.put(dout.data(), ng);
  smoothed_map(dmr.data(), dout.data(), L, dmt.data(), thr_d);  // response, variable resolution
  F.put_int(nb);
  F.put(dout.data(), ng);
  F.put_int(nb);  // response, fine resolution
  F.put(&dmr[GALIC_STACKOFFSET(L, 0, 0)], ng);
  // cell-centre coordinates of the density and the energy grid (io.c:292-361)
  std::vector<double> R(ng), z(ng);
  const double rmin[2] = {p.DG_Rmin, p.EG_Rmin}, fac[2] = {p.DG_Fac, p.EG_Fac};
  for(int g = 0; g < 2; g++) {
    for(int k = 0; k < nb; k++) {
      const double zc = rmin[g] * 0.5 * (pow(fac[g], k) + pow(fac[g], k + 1) - 2.0);
      for(int j = 0; j < nb; j++) {
        R[k * nb + j] = rmin[g] * 0.5 * (pow(fac[g], j) + pow(fac[g], j + 1) - 2.0);
        z[k * nb + j] = zc;
      }
    }
    F.put_int(nb);
    F.put(R.data(), ng);
    F.put_int(nb);
    F.put(z.data(), ng);
  }
  // energy grid: mass target and response, then each moment's target and response,
  // all at the resolution the mass response supports (io.c:367-452)
  const std::vector<double> *pairs[5][2] = {{&emt, &emr}, {&et[0], &er[0]}, {&et[1], &er[1]}, {&et[2], &er[2]},
                                            {&et[3], &er[3]}};
  for(int k = 0; k < 5; k++)
    for(int h = 0; h < 2; h++) {
      smoothed_map(pairs[k][h]->data(), dout.data(), L, emr.data(), thr_s);
      F.put_int(nb);
      F.put(dout.data(), ng);
    }
  if(!F.ok) return io_err(GALIC_ERR_ARG, "writing the densfield file failed");
  const int cl = fclose(F.f);
  F.f = nullptr;
  return cl == 0 ? GALIC_OK : io_err(GALIC_ERR_ARG, "closing the densfield file failed");
}

int galic_write_snapshot(galic_ctx *ctx, const char *path, int double_precision, const double *vel_theo) {
  if(!ctx || !path) return io_err(GALIC_ERR_ARG, "bad arguments");
  galic::Ctx &c = ctx->c;
  const int n = c.P.n;
  if(n <= 0) return io_err(GALIC_ERR_STATE, "no particles");
  std::vector<double> x(n), y(n), zz(n), vx(n), vy(n), vz(n), m(n);
  std::vector<int> t(n);
  IO_HIP(hipSetDevice(c.device));
  IO_HIP(hipStreamSynchronize(c.stream));
  double *const src[7] = {c.P.x, c.P.y, c.P.z, c.P.vx, c.P.vy, c.P.vz, c.P.mass};
  std::vector<double> *dst[7] = {&x, &y, &zz, &vx, &vy, &vz, &m};
  for(int k = 0; k < 7; k++) IO_HIP(hipMemcpy(dst[k]->data(), src[k], n * sizeof(double), hipMemcpyDeviceToHost));
  IO_HIP(hipMemcpy(t.data(), c.P.type, n * sizeof(int), hipMemcpyDeviceToHost));
  // particles type by type, in their order within a type (fill_write_buffer, io.c:553-600)
  std::vector<int> order;
  order.reserve(n);
  int npart[6] = {0, 0, 0, 0, 0, 0};
  for(int ty = 0; ty < 6; ty++)
    for(int i = 0; i < n; i++)
      if(t[i] == ty) {
        order.push_back(i);
        npart[ty]++;
      }
  // header (allvars.h:998-1031), filled as write_file does (io.c:952-985); all masses are
  // stored per particle (MassTable is zero)
  unsigned char hdr[256];
  memset(hdr, 0, sizeof(hdr));
  size_t o = 0;
  auto h_put = [&](const void *v, size_t b) {
    memcpy(hdr + o, v, b);
    o += b;
  };
  h_put(npart, sizeof(npart));
  const double zero6[6] = {0, 0, 0, 0, 0, 0};
  h_put(zero6, sizeof(zero6));  // mass[6]
  const double zero = 0.0;
  h_put(&zero, 8);  // time
  h_put(&zero, 8);  // redshift
  const int izero = 0, one = 1;
  h_put(&izero, 4);  // flag_sfr
  h_put(&izero, 4);  // flag_feedback
  unsigned int tot[6];
  for(int k = 0; k < 6; k++) tot[k] = (unsigned int)npart[k];
  h_put(tot, sizeof(tot));  // npartTotal
  h_put(&izero, 4);         // flag_cooling
  h_put(&one, 4);           // num_files
  for(int k = 0; k < 4; k++) h_put(&zero, 8);  // BoxSize, Omega0, OmegaLambda, HubbleParam
  h_put(&izero, 4);                             // flag_stellarage
  h_put(&izero, 4);                             // flag_metals
  const unsigned int zero6u[6] = {0, 0, 0, 0, 0, 0};
  h_put(zero6u, sizeof(zero6u));  // npartTotalHighWord
  h_put(&izero, 4);               // flag_entropy_instead_u
  const int dp = double_precision ? 1 : 0;
  h_put(&dp, 4);  // flag_doubleprecision
  // flag_lpt_ics, lpt_scalingfactor, flag_tracer_field, composition_vector_length and the fill stay zero
  File F;
  F.f = fopen(path, "wb");
  if(!F.f) return io_err(GALIC_ERR_ARG, "cannot create the snapshot file");
  auto block = [&](int bytes) { F.put_int(bytes); };  // the record length before and after a block
  block(256);
  F.put(hdr, 256);
  block(256);
  const int fsz = dp ? 8 : 4;
  auto vec3 = [&](const std::vector<double> &a, const std::vector<double> &b, const std::vector<double> &cc) {
    block(3 * fsz * n);
    for(int i : order) {
      const double v[3] = {a[i], b[i], cc[i]};
      if(dp) {
        F.put(v, 3);
      } else {
        const float f3[3] = {(float)v[0], (float)v[1], (float)v[2]};
        F.put(f3, 3);
      }
    }
    block(3 * fsz * n);
  };
  vec3(x, y, zz);   // IO_POS
  vec3(vx, vy, vz); // IO_VEL
  block(4 * n);     // IO_ID: set_particles.c:65 numbers the particles 1..N in their order
  for(int i : order) {
    const unsigned int id = (unsigned int)(i + 1);
    F.put(&id, 1);
  }
  block(4 * n);
  block(fsz * n);  // IO_MASS
  for(int i : order) {
    if(dp) {
      F.put(&m[i], 1);
    } else {
      const float f = (float)m[i];
      F.put(&f, 1);
    }
  }
  block(fsz * n);
  // IO_VELTHEO: the comparison velocities of set_particles.c:94-109, zero if not given
  std::vector<double> tx(n, 0.0), ty(n, 0.0), tz(n, 0.0);
  if(vel_theo)
    for(int i = 0; i < n; i++) {
      tx[i] = vel_theo[3 * i];
      ty[i] = vel_theo[3 * i + 1];
      tz[i] = vel_theo[3 * i + 2];
    }
  vec3(tx, ty, tz);
  if(!F.ok) return io_err(GALIC_ERR_ARG, "writing the snapshot failed");
  const int cl = fclose(F.f);
  F.f = nullptr;
  return cl == 0 ? GALIC_OK : io_err(GALIC_ERR_ARG, "closing the snapshot failed");
}

}  // extern "C"
