// galic_targets.hip -- sampled density and dispersion targets on the device.
//
// densitygrid_sample_targetresponse (grid.c:536-617) and
// calc_energy_grid_mass_maps (grid.c:941-1076) draw fresh points of one
// component (halo.c:17-40, disk.c:17-51, bulge.c:17-38), keep those with
// sqrt(x^2 + z^2) < Rmax, and add each point's four cloud-in-cell weights to a
// mass map on the density grid, and the weights times the point's target
// dispersions (get_disp_rtp, disp_fields.c:66-115) to four maps on the energy
// grid.  One thread takes sample ids in turn; a sample's random numbers are
// Philox draws of its id (attempt = the rejection round), so the points do not
// depend on the launch.  The maps are summed with 64-bit integer atomics on
// fixed-point weights, so they are the same bits whatever the order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "galic_rng.h"
#include "galic_targets.h"

namespace galic {

namespace {

constexpr int kMaxAttempts = 1000;  // rejection rounds per sample (one round is rejected with p < 0.1 for a
                                    // grid holding most of the mass)
constexpr int kMaxNewton = 200;     // Newton iterations of the disk radius (converges in < 10 from R = 1)

// A fresh point of the component, from the Philox numbers of (id, attempt).
// Returns false if no round was accepted.
__device__ bool sample_point(const TargetArgs &A, uint64_t id, double pos[3]) {
  for(uint32_t attempt = 0; attempt < (uint32_t)kMaxAttempts; attempt++) {
    double u[4];
    uniform4(id, attempt, 0, A.seed, A.stream, u);
    if(A.type == 2) {
      // disk.c:17-51: sech^2 height by inversion, exponential radius by Newton
      // iteration on (1 + R) exp(-R) = 1 - q
      const double z = A.Disk_Z0 / 2 * log(u[0] / (1 - u[0]));
      const double q = u[1];
      double R = 1.0;
      for(int it = 0; it < kMaxNewton; it++) {
        const double f = (1 + R) * exp(-R) + q - 1;
        const double f_ = -R * exp(-R);
        const double Rold = R;
        R = R - f / f_;
        if(!(fabs(R - Rold) / R > 1e-7)) break;
      }
      R *= A.Disk_H;
      const double phi = u[2] * M_PI * 2;
      pos[0] = R * cos(phi);
      pos[1] = R * sin(phi);
      pos[2] = z;
      if(!(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2] > A.Rmax * A.Rmax)) return true;
    } else {
      // halo.c:17-40 and bulge.c:17-38: Hernquist radius by inversion of
      // M(<r)/M = r^2/(r+a)^2, isotropic direction, z divided by the stretch
      const double a = A.type == 1 ? A.M.Halo_A : A.M.Bulge_A;
      const double q = u[0];
      const double r = q > 0 ? a * (q + sqrt(q)) / (1 - q) : 0.0;
      if(A.type == 3 && r > A.Rmax) continue;  // the bulge tests the radius before the angles
      const double phi = u[1] * M_PI * 2;
      const double theta = acos(u[2] * 2 - 1);
      pos[0] = r * sin(theta) * cos(phi);
      pos[1] = r * sin(theta) * sin(phi);
      pos[2] = r * cos(theta) / A.stretch;
      if(A.type == 3) return true;
      if(!(sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]) > A.Rmax)) return true;
    }
  }
  return false;
}

// Bilinear value of a force-grid table through forcegrid_get_cell (grid.c:15-55).
__device__ __forceinline__ double fg_bilinear(const DevModel &M, const FgIndex64 &g, const double *T) {
  const int nb = M.FG_Nbin;
  double br = g.jdbl, bz = g.idbl;
  if(br < 0) br = 0;
  if(bz < 0) bz = 0;
  int iR = (int)br;
  double fR = br - iR;
  int iz = (int)bz;
  double fz = bz - iz;
  if(iR >= nb - 1) {
    iR = nb - 2;
    fR = 1;
  }
  if(iz >= nb - 1) {
    iz = nb - 2;
    fz = 1;
  }
  return T[iz * nb + iR] * (1 - fR) * (1 - fz) + T[(iz + 1) * nb + iR] * (1 - fR) * (fz) +
         T[iz * nb + (iR + 1)] * (fR) * (1 - fz) + T[(iz + 1) * nb + (iR + 1)] * (fR) * (fz);
}

// get_radial_disp_spherical (disp_fields.c:375-398): linear in the log radial index.
__device__ __forceinline__ double radial_disp(const TargetArgs &A, const double pos[3]) {
  const int nb = A.M.FG_Nbin;
  const double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
  double br = (log(r / A.M.FG_Rmin + 1.0) / A.M.FG_logFac);
  if(br < 0) br = 0;
  int binr = (int)br;
  double fR = br - binr;
  if(binr >= nb - 1) {
    binr = nb - 2;
    fR = 1;
  }
  return A.disp_r[binr] * (1 - fR) + A.disp_r[binr + 1] * (fR);
}

// get_beta_of_type (disp_fields.c:26-64)
__device__ __forceinline__ double beta_of_type(const TargetArgs &A, const double pos[3]) {
  if(A.type == 2) return 0.0;
  double beta = A.beta;
  if(beta >= 1) {
    const double a = A.type == 1 ? A.M.Halo_A : A.M.Bulge_A;
    const double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
    const double dlogrhodr = -1.0 - 3.0 * r / (r + a);
    beta = -0.15 - 0.20 * dlogrhodr;
  }
  return beta;
}

// get_disp_rtp (disp_fields.c:66-115)
__device__ __forceinline__ void disp_rtp(const TargetArgs &A, const double pos[3], double d[4]) {
  if(A.tovs == 0) {
    d[0] = radial_disp(A, pos);
    d[1] = d[2] = d[3] = d[0];
  } else if(A.tovs == 1) {
    d[0] = radial_disp(A, pos);
    d[1] = (1 - beta_of_type(A, pos)) * d[0];
    d[2] = d[3] = d[1];
  } else {
    const FgIndex64 g = fg_index64(A.M, pos[0], pos[1], pos[2]);
    if(A.tovs == 2) {
      d[1] = fg_bilinear(A.M, g, A.disp_z);
      d[0] = d[1];
    } else {
      d[0] = fg_bilinear(A.M, g, A.tilt_r);
      d[1] = fg_bilinear(A.M, g, A.tilt_theta);
    }
    d[2] = fg_bilinear(A.M, g, A.disp_phi);
    const double vstr = vstream64(A.M, g, A.type);
    d[3] = d[2] + vstr * vstr;
  }
}

__device__ __forceinline__ unsigned long long to_q(double v) { return (unsigned long long)__double2ll_rn(v); }

// The maps of one launch (A.dg and/or A.eg maps f0 .. f1-1), summed per
// workgroup in LDS when A.lds and added to memory at the end, else added to
// memory directly.  Integer sums: the result is the same bits either way.
__global__ void __launch_bounds__(256) k_sample_targets(TargetArgs A) {
  extern __shared__ unsigned long long smap[];
  const int ngd = A.M.DG_Nbin * A.M.DG_Nbin, nge = A.M.EG_Nbin * A.M.EG_Nbin;
  const int ndl = A.dg ? ngd : 0, nl = ndl + (A.eg ? (A.f1 - A.f0) * nge : 0);
  unsigned long long *sdg = smap, *seg = smap + ndl;
  if(A.lds) {
    for(int q = threadIdx.x; q < nl; q += blockDim.x) smap[q] = 0;
    __syncthreads();
  }
  const long long stride = (long long)gridDim.x * blockDim.x;
  unsigned long long binned = 0, failed = 0;
  for(long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < A.n; id += stride) {
    double pos[3] = {0, 0, 0};
    const bool ok = sample_point(A, (uint64_t)id, pos);
    if(A.pos_out) {
      A.pos_out[3 * id] = ok ? pos[0] : 0.0;
      A.pos_out[3 * id + 1] = ok ? pos[1] : 0.0;
      A.pos_out[3 * id + 2] = ok ? pos[2] : 0.0;
    }
    if(!ok) {
      failed++;
      continue;
    }
    // the reference keeps points with sqrt(x^2 + z^2) < Rmax (grid.c:568-570, 985-987)
    if(!(sqrt(pos[0] * pos[0] + pos[2] * pos[2]) < A.Rmax)) continue;
    binned++;
    if(A.dg) {
      int iR, iz;
      double fR, fz;
      cell64(pos[0], pos[1], pos[2], A.M.DG_Rmin, A.M.DG_logFac, A.M.DG_Nbin, iR, iz, fR, fz);
      const int nb = A.M.DG_Nbin;
      const double w[4] = {(1 - fR) * (1 - fz), (fR) * (1 - fz), (1 - fR) * (fz), (fR) * (fz)};
      const int c[4] = {iz * nb + iR, iz * nb + (iR + 1), (iz + 1) * nb + iR, (iz + 1) * nb + (iR + 1)};
      for(int k = 0; k < 4; k++) {
        const unsigned long long q = to_q(A.qdg * w[k]);
        if(A.lds)
          atomicAdd(&sdg[c[k]], q);
        else
          atomicAdd((unsigned long long *)&A.dg[c[k]], q);
      }
    }
    if(A.eg && A.f1 > A.f0) {
      int iR, iz;
      double fR, fz;
      cell64(pos[0], pos[1], pos[2], A.M.EG_Rmin, A.M.EG_logFac, A.M.EG_Nbin, iR, iz, fR, fz);
      const int nb = A.M.EG_Nbin;
      double d[4] = {0, 0, 0, 0};
      if(A.f1 > 1) disp_rtp(A, pos, d);
      const double w[4] = {(1 - fR) * (1 - fz), (fR) * (1 - fz), (1 - fR) * (fz), (fR) * (fz)};
      const int c[4] = {iz * nb + iR, iz * nb + (iR + 1), (iz + 1) * nb + iR, (iz + 1) * nb + (iR + 1)};
      for(int k = 0; k < 4; k++)
        for(int f = A.f0; f < A.f1; f++) {
          const unsigned long long q = to_q(A.qeg[f] * (f == 0 ? w[k] : w[k] * d[f - 1]));
          if(A.lds)
            atomicAdd(&seg[(size_t)(f - A.f0) * nge + c[k]], q);
          else
            atomicAdd((unsigned long long *)&A.eg[(size_t)f * nge + c[k]], q);
        }
    }
  }
  if(A.lds) {
    __syncthreads();
    for(int q = threadIdx.x; q < nl; q += blockDim.x) {
      const unsigned long long v = smap[q];
      if(v == 0) continue;
      unsigned long long *dst = q < ndl ? (unsigned long long *)&A.dg[q]
                                        : (unsigned long long *)&A.eg[(size_t)A.f0 * nge + (q - ndl)];
      atomicAdd(dst, v);
    }
  }
  if(A.count) {
    if(binned) atomicAdd(&A.counts[0], binned);
    if(failed) atomicAdd(&A.counts[1], failed);
  }
}

}  // namespace

int launch_sample_targets(const TargetArgs &A, int grid, size_t shmem, hipStream_t s) {
  hipLaunchKernelGGL(k_sample_targets, dim3(grid), dim3(256), shmem, s, A);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace galic
