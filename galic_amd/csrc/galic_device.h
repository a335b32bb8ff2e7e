// galic_device.h -- device-side model description and field lookups for gfx950.
//
// The FP64 lookups keep the reference's operation order (divisions stay
// divisions, log(Fac) is divided, not multiplied by a reciprocal) so the plain
// path reproduces the reference to rounding:
//   forcegrid_get_cell        grid.c:15-55
//   forcegrid_get_potential   grid.c:59-113
//   forcegrid_get_acceleration grid.c:122-197
//   densitygrid_get_cell      grid.c:494-534
//   energygrid_get_cell       grid.c:896-938
//   get_vstream               disp_fields.c:118-132
//   halo/bulge closed forms   halo.c:63-84, bulge.c:64-85
// Two exact rewrites are used: log(FG_Fac) and log(DG_Fac) are evaluated once
// (the reference recomputes the same value at every call), and the force-grid
// log index is shared between the acceleration and the streaming-velocity
// lookups at the same position (identical expressions, identical results).
// cos(atan2(y,x)) and sin(atan2(y,x)) become x/R and y/R, which differ from the
// reference by at most a couple of ulp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/galic_types.h"

namespace galic {

constexpr int kFields = 5;  // mass, r, t, q, p (orbit_response.c:155-157 argument order)

// Scalars and table pointers a kernel needs; passed by value.
struct DevModel {
  double G, V200, R200, M200;
  double Halo_Mass, Halo_A, Bulge_Mass, Bulge_A;
  double halo_offgrid_fac;  // M200 / Halo_Mass (grid.c:86,160)
  int use_table;            // grid.c:63,124
  int add_halo, add_bulge;  // grid.c:94,103,172,185
  int tovs[GALIC_NTYPES];
  double fo, fc;            // TimeStepFactorOrbit, TimeStepFactorCellCross
  long long max_steps;

  int FG_Nbin;
  double FG_Rmin, FG_logFac;
  int DG_Nbin, DG_L;
  double DG_Rmin, DG_logFac;
  int EG_Nbin, EG_L;
  double EG_Rmin, EG_logFac;

  // FP32 forms of the grid constants (reciprocals are fine in the FP32 path)
  float fFG_invRmin, fFG_invlogFac, fDG_invRmin, fDG_invlogFac;

  const double *FG_Pot, *FG_DPotDR, *FG_DPotDz;
  const double *FG_Vstream[GALIC_NTYPES];
  const double *DG_CellSize;
  // FP32 tables: accR/accz interleaved as float2 per node, vstream per type
  const float2 *FG_Acc_f;
  const float *FG_Vstream_f[GALIC_NTYPES];
  const float *DG_CellSize_f;
};

// ---------------------------------------------------------------- FP64 ----

__device__ __forceinline__ double halo_pot64(const DevModel &M, double x, double y, double z) {
  double r = sqrt(x * x + y * y + z * z);
  return -M.G * M.Halo_Mass / (r + M.Halo_A);
}

__device__ __forceinline__ double halo_accfac64(const DevModel &M, double x, double y, double z) {
  double r = sqrt(x * x + y * y + z * z);
  return M.G * M.Halo_Mass / ((r + 1.0e-6 * M.Halo_A) * (r + M.Halo_A) * (r + M.Halo_A));
}

__device__ __forceinline__ double bulge_pot64(const DevModel &M, double x, double y, double z) {
  double r = sqrt(x * x + y * y + z * z);
  return -M.G * M.Bulge_Mass / (r + M.Bulge_A);
}

__device__ __forceinline__ double bulge_accfac64(const DevModel &M, double x, double y, double z) {
  double r = sqrt(x * x + y * y + z * z);
  return M.G * M.Bulge_Mass / ((r + 1.0e-6 * M.Bulge_A) * (r + M.Bulge_A) * (r + M.Bulge_A));
}

// Log index of the force grid at (R,|z|): jdbl = log(R/Rmin+1)/log(Fac) (grid.c:20-21, 130-134)
struct FgIndex64 {
  double R, jdbl, idbl;
};

__device__ __forceinline__ FgIndex64 fg_index64(const DevModel &M, double x, double y, double z) {
  FgIndex64 g;
  g.R = sqrt(x * x + y * y);
  g.jdbl = log(g.R / M.FG_Rmin + 1.0) / M.FG_logFac;
  g.idbl = log(fabs(z) / M.FG_Rmin + 1.0) / M.FG_logFac;
  return g;
}

// forcegrid_get_acceleration, grid.c:122-197
__device__ __forceinline__ void accel64(const DevModel &M, double x, double y, double z, const FgIndex64 &g,
                                        double &ax, double &ay, double &az) {
  ax = ay = az = 0;
  if(M.use_table) {
    const int nb = M.FG_Nbin;
    int jbin = (int)g.jdbl;
    double jfac = g.jdbl - jbin;
    int ibin = (int)g.idbl;
    double ifac = g.idbl - ibin;
    if(ibin < nb - 1 && jbin < nb - 1) {
      const double *T = M.FG_DPotDR;
      const double *U = M.FG_DPotDz;
      int o = ibin * nb + jbin;
      double accR = T[o] * (1 - ifac) * (1 - jfac) + T[o + 1] * (1 - ifac) * jfac + T[o + nb] * ifac * (1 - jfac) +
                    T[o + nb + 1] * ifac * jfac;
      double accz = U[o] * (1 - ifac) * (1 - jfac) + U[o + 1] * (1 - ifac) * jfac + U[o + nb] * ifac * (1 - jfac) +
                    U[o + nb + 1] * ifac * jfac;
      if(z < 0) accz = -accz;
      double c = 1.0, s = 0.0;
      if(g.R > 0) {
        c = x / g.R;
        s = y / g.R;
      }
      ax = accR * c;
      ay = accR * s;
      az = accz;
    } else {
      double fac = halo_accfac64(M, x, y, z);
      ax = -fac * x * M.halo_offgrid_fac;
      ay = -fac * y * M.halo_offgrid_fac;
      az = -fac * z * M.halo_offgrid_fac;
    }
  }
  if(M.add_halo) {
    double fac = halo_accfac64(M, x, y, z);
    ax += -fac * x;
    ay += -fac * y;
    az += -fac * z;
  }
  if(M.add_bulge) {
    double fac = bulge_accfac64(M, x, y, z);
    ax += -fac * x;
    ay += -fac * y;
    az += -fac * z;
  }
}

// forcegrid_get_potential, grid.c:59-113
__device__ __forceinline__ double potential64(const DevModel &M, double x, double y, double z) {
  double pot = 0;
  if(M.use_table) {
    FgIndex64 g = fg_index64(M, x, y, z);
    const int nb = M.FG_Nbin;
    int jbin = (int)g.jdbl;
    double jfac = g.jdbl - jbin;
    int ibin = (int)g.idbl;
    double ifac = g.idbl - ibin;
    if(ibin < nb - 1 && jbin < nb - 1) {
      const double *P = M.FG_Pot;
      int o = ibin * nb + jbin;
      pot = P[o] * (1 - ifac) * (1 - jfac) + P[o + 1] * (1 - ifac) * jfac + P[o + nb] * ifac * (1 - jfac) +
            P[o + nb + 1] * ifac * jfac;
    } else {
      pot = M.halo_offgrid_fac * halo_pot64(M, x, y, z);
    }
  }
  if(M.add_halo) pot += halo_pot64(M, x, y, z);
  if(M.add_bulge) pot += bulge_pot64(M, x, y, z);
  return pot;
}

// get_vstream (disp_fields.c:118-132) through forcegrid_get_cell (grid.c:15-55)
__device__ __forceinline__ double vstream64(const DevModel &M, const FgIndex64 &g, int type) {
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
  const double *V = M.FG_Vstream[type];
  return V[iz * nb + iR] * (1 - fR) * (1 - fz) + V[(iz + 1) * nb + iR] * (1 - fR) * (fz) +
         V[iz * nb + (iR + 1)] * (fR) * (1 - fz) + V[(iz + 1) * nb + (iR + 1)] * (fR) * (fz);
}

// densitygrid_get_cell (grid.c:494-534); with EG constants it is energygrid_get_cell (grid.c:896-938)
__device__ __forceinline__ void cell64(double x, double y, double z, double Rmin, double logFac, int nbin, int &iR,
                                       int &iz, double &fR, double &fz) {
  double r = sqrt(x * x + y * y);
  double az = fabs(z);
  double br = (log(r / Rmin + 1.0) / logFac) - 0.5;
  double bz = (log(az / Rmin + 1.0) / logFac) - 0.5;
  if(br < 0) br = 0;
  if(bz < 0) bz = 0;
  iR = (int)br;
  fR = br - iR;
  iz = (int)bz;
  fz = bz - iz;
  if(iR >= nbin - 1) {
    iR = nbin - 2;
    fR = 1;
  }
  if(iz >= nbin - 1) {
    iz = nbin - 2;
    fz = 1;
  }
}

// Velocity decomposition of orbit_response.c:222-273 (cylindrical basis for
// velocity structure 2, spherical otherwise).
__device__ __forceinline__ void decompose64(const DevModel &M, int tovs, int type, double x, double y, double z,
                                            double vx, double vy, double vz, const FgIndex64 &g, double &vr2,
                                            double &vt2, double &vq2, double &vp2) {
  if(tovs == 2) {
    double r2 = x * x + y * y;
    double v_dot_x = vx * x + vy * y;
    vr2 = v_dot_x * v_dot_x / r2;
    double Q0 = -y, Q1 = x;
    double q2 = Q0 * Q0 + Q1 * Q1;
    double v_dot_Q = vx * Q0 + vy * Q1;
    double q = sqrt(q2);
    double vq = v_dot_Q / q;
    vq2 = vq * vq;
    double vstr = vstream64(M, g, type);
    vp2 = (vq - vstr) * (vq - vstr);
    vt2 = vz * vz;
  } else {
    // Z = (0,0,-1); Q = x cross Z; T = x cross Q (orbit_response.c:247-272)
    double r2 = x * x + y * y + z * z;
    double v_dot_x = vx * x + vy * y + vz * z;
    vr2 = v_dot_x * v_dot_x / r2;
    double Q0 = y * (-1.0) - z * 0.0;
    double Q1 = z * 0.0 - x * (-1.0);
    double Q2 = x * 0.0 - y * 0.0;
    double q2 = Q0 * Q0 + Q1 * Q1 + Q2 * Q2;
    double v_dot_Q = vx * Q0 + vy * Q1 + vz * Q2;
    double q = sqrt(q2);
    double vq = v_dot_Q / q;
    vq2 = vq * vq;
    double vstr = vstream64(M, g, type);
    vp2 = (vq - vstr) * (vq - vstr);
    double T0 = y * Q2 - z * Q1;
    double T1 = z * Q0 - x * Q2;
    double T2 = x * Q1 - y * Q0;
    double t2 = T0 * T0 + T1 * T1 + T2 * T2;
    double v_dot_T = vx * T0 + vy * T1 + vz * T2;
    vt2 = v_dot_T * v_dot_T / t2;
  }
}

// Instantaneous squared velocity components at the particle position,
// calc_disp_components_for_particle (disp_fields.c:200-372).
__device__ __forceinline__ void disp_components64(const DevModel &M, int type, double x, double y, double z,
                                                  double vx, double vy, double vz, double &vr2, double &vt2,
                                                  double &vp2, double &vq2) {
  int tovs = M.tovs[type];
  if(tovs == 0 || tovs == 1 || tovs == 3) {
    double phi = atan2(y, x);
    double theta = acos(z / sqrt(x * x + y * y + z * z));
    double st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
    double er0 = st * cp, er1 = st * sp, er2 = ct;
    double ep0 = -sp, ep1 = cp, ep2 = 0;
    double et0 = -ct * cp, et1 = -ct * sp, et2 = st;
    double vr = vx * er0 + vy * er1 + vz * er2;
    double vphi = vx * ep0 + vy * ep1 + vz * ep2;
    double vtheta = vx * et0 + vy * et1 + vz * et2;
    double vstr = 0;
    if(tovs == 1 || tovs == 3) vstr = vstream64(M, fg_index64(M, x, y, z), type);
    vr2 = vr * vr;
    vt2 = vtheta * vtheta;
    vp2 = (vphi - vstr) * (vphi - vstr);
    vq2 = vphi * vphi;
  } else {
    double phi = atan2(y, x);
    double sp = sin(phi), cp = cos(phi);
    double vR = vx * cp + vy * sp + vz * 0;
    double vphi = vx * (-sp) + vy * cp + vz * 0;
    double vZ = vx * 0 + vy * 0 + vz * 1;
    double vstr = vstream64(M, fg_index64(M, x, y, z), type);
    vr2 = vR * vR;
    vt2 = vZ * vZ;
    vp2 = (vphi - vstr) * (vphi - vstr);
    vq2 = vphi * vphi;
  }
}

}  // namespace galic
