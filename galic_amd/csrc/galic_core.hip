// galic_core.hip -- the plain FP64 path on gfx950: orbit integration with
// binning, the incremental merit with accept/reject, commit and global fit.
//
// Built with -ffp-contract=off: every a*b+c rounds twice, as in the reference's
// x86-64 build, so results agree with the reference to rounding.
//
// Kernels and the reference code they replace:
//   k_orbit64        produce_orbit_response_field_mod   orbit_response.c:155-491
//   k_inst_init      calc_disp_components + add_to_energy_grid in
//                    calc_all_response_fields           main.c:226-229
//   k_eval_apply64   eval_fit_mod x2 + accept + deltas  main.c:460-505, 1021-1217
//   k_commit         commit_updates                     main.c:808-841
//   k_fit_stacks     calc_global_fit: Allreduce+smooth  main.c:256-292
//   k_fit_sums       calc_global_fit: merit sums        main.c:294-359
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "galic_ctx.h"
#include "galic_kernels.h"
#include "galic_merit.h"

namespace galic {

// ------------------------------------------------------------------ orbits --

// Register-resident accumulator for consecutive binning events that hit the
// same 2x2 block of cells: flushed to memory only when the block changes.
struct Run64 {
  int quad;  // iz*Nbin + iR of the lower-left cell, -1 when empty
  double w[kFields][4];
};

__device__ __forceinline__ void run_reset(Run64 &r, int quad) {
  r.quad = quad;
#pragma unroll
  for(int f = 0; f < kFields; f++)
#pragma unroll
    for(int c = 0; c < 4; c++) r.w[f][c] = 0;
}

// Write a run either into the orbit's private slot (trial evaluation) or with
// atomics into the Loc grid of the particle's type (initial response pass).
__device__ __forceinline__ void run_flush(const Run64 &r, int nb, int ngrid, double fac, double *slot,
                                          double *loc) {
  if(r.quad < 0) return;
  const int c0 = r.quad, c1 = r.quad + 1, c2 = r.quad + nb, c3 = r.quad + nb + 1;
  if(slot) {
    // cell-major slot: the 5 fields of a cell are contiguous, so a flush
    // touches two short row segments instead of ten field rows
    double *s0 = slot + (size_t)c0 * kFields, *s2 = slot + (size_t)c2 * kFields;
#pragma unroll
    for(int f = 0; f < kFields; f++) {
      s0[f] += fac * r.w[f][0];
      s0[kFields + f] += fac * r.w[f][1];
      s2[f] += fac * r.w[f][2];
      s2[kFields + f] += fac * r.w[f][3];
    }
  } else {
#pragma unroll
    for(int f = 0; f < kFields; f++) {
      double *s = loc + (size_t)f * ngrid;
      atomicAdd(&s[c0], fac * r.w[f][0]);
      atomicAdd(&s[c1], fac * r.w[f][1]);
      atomicAdd(&s[c2], fac * r.w[f][2]);
      atomicAdd(&s[c3], fac * r.w[f][3]);
    }
  }
}

// One binning event: trapezoid half weights (orbit_response.c:277-319).
__device__ __forceinline__ void bin_event64(Run64 &r, int nb, int ngrid, double fac, double *slot, double *loc,
                                            int iR, int iz, double fR, double fz, double dt, double vr2,
                                            double vt2, double vq2, double vp2, int4 &bb) {
  double m00 = 0.5 * dt * (1 - fR) * (1 - fz);
  double m10 = 0.5 * dt * (fR) * (1 - fz);
  double m01 = 0.5 * dt * (1 - fR) * (fz);
  double m11 = 0.5 * dt * (fR) * (fz);
  int quad = iz * nb + iR;
  if(quad != r.quad) {
    run_flush(r, nb, ngrid, fac, slot, loc);
    run_reset(r, quad);
    bb.x = min(bb.x, iR);
    bb.y = max(bb.y, iR + 1);
    bb.z = min(bb.z, iz);
    bb.w = max(bb.w, iz + 1);
  }
  r.w[F_M][0] += m00; r.w[F_M][1] += m10; r.w[F_M][2] += m01; r.w[F_M][3] += m11;
  r.w[F_R][0] += m00 * vr2; r.w[F_R][1] += m10 * vr2; r.w[F_R][2] += m01 * vr2; r.w[F_R][3] += m11 * vr2;
  r.w[F_T][0] += m00 * vt2; r.w[F_T][1] += m10 * vt2; r.w[F_T][2] += m01 * vt2; r.w[F_T][3] += m11 * vt2;
  r.w[F_Q][0] += m00 * vq2; r.w[F_Q][1] += m10 * vq2; r.w[F_Q][2] += m01 * vq2; r.w[F_Q][3] += m11 * vq2;
  r.w[F_P][0] += m00 * vp2; r.w[F_P][1] += m10 * vp2; r.w[F_P][2] += m01 * vp2; r.w[F_P][3] += m11 * vp2;
}

// Persistent lanes: every lane pulls the next orbit from an atomic counter when
// its current orbit ends, so orbits of very different length keep the wave busy.
// Every lane leaves once the counter passes the orbit count.
__global__ void __launch_bounds__(256) k_orbit64(DevModel M, OrbitArgs A) {
  const int nb = M.DG_Nbin, ngrid = nb * nb;
  while(true) {
    int o = (int)atomicAdd(A.counter, 1u);
    if(o >= A.n) break;
    const int type = A.otype[o];
    const int tovs = M.tovs[type];
    double x = A.ox[o], y = A.oy[o], z = A.oz[o];
    double vx = A.ovx[o], vy = A.ovy[o], vz = A.ovz[o];
    const double timespan = A.otint[o];
    const double fac = A.omass[o] / timespan;  // mass/t with t == timespan at the end (see tall)
    double *slot = A.slots ? A.slots + (size_t)(o - A.slot_base) * A.slot_stride : nullptr;
    double *loc = A.loc ? A.loc + (size_t)type * kFields * ngrid : nullptr;

    FgIndex64 g = fg_index64(M, x, y, z);
    double ax, ay, az;
    accel64(M, x, y, z, g, ax, ay, az);
    int iR, iz;
    double fR, fz;
    cell64(x, y, z, M.DG_Rmin, M.DG_logFac, nb, iR, iz, fR, fz);
    int icell = iz * nb + iR;
    const double E0 = 0.5 * (vx * vx + vy * vy + vz * vz) + potential64(M, x, y, z);
    double vr2, vt2, vq2, vp2;
    decompose64(M, tovs, type, x, y, z, vx, vy, vz, g, vr2, vt2, vq2, vp2);

    Run64 run;
    run.quad = -1;
    int4 bb = make_int4(nb, -1, nb, -1);
    double tall = 0, radsign_previous = 0;
    int norbit = 0, flag = 0, status = 0;
    long long steps = 0;
    if(!(timespan > 0)) status = GALIC_ERR_ARG;  // the reference would loop until its step cap
    while(!flag && status == 0) {
      // get_timestep (orbit_response.c:16-26), with the cell of the previous position
      double v = sqrt(vx * vx + vy * vy + vz * vz);
      double aa = sqrt(ax * ax + ay * ay + az * az);
      double torbit = M.V200 / aa;
      double tcross = M.DG_CellSize[icell] / v;
      double t1 = M.fo * torbit, t2 = M.fc * tcross;
      double dt = (t1 < t2) ? t1 : t2;
      if(timespan < dt + tall) {
        dt = timespan - tall;
        flag = 1;
      }
      if(!(dt >= 0 && dt <= timespan)) {  // NaN state: stop instead of running to the step cap
        status = GALIC_ERR_ARG;
        break;
      }
      bin_event64(run, nb, ngrid, fac, slot, loc, iR, iz, fR, fz, dt, vr2, vt2, vq2, vp2, bb);
      vx += 0.5 * dt * ax;
      vy += 0.5 * dt * ay;
      vz += 0.5 * dt * az;
      x += dt * vx;
      y += dt * vy;
      z += dt * vz;
      g = fg_index64(M, x, y, z);
      accel64(M, x, y, z, g, ax, ay, az);
      vx += 0.5 * dt * ax;
      vy += 0.5 * dt * ay;
      vz += 0.5 * dt * az;
      cell64(x, y, z, M.DG_Rmin, M.DG_logFac, nb, iR, iz, fR, fz);
      icell = iz * nb + iR;
      // the decomposition at the end of this step is also the one at the start of the next
      decompose64(M, tovs, type, x, y, z, vx, vy, vz, g, vr2, vt2, vq2, vp2);
      bin_event64(run, nb, ngrid, fac, slot, loc, iR, iz, fR, fz, dt, vr2, vt2, vq2, vp2, bb);
      tall += dt;
      double radsign = vx * x + vy * y + vz * z;
      if(radsign > 0 && radsign_previous < 0) norbit++;
      radsign_previous = radsign;
      steps++;
      if(steps > M.max_steps) status = GALIC_ERR_STEPCAP;
    }
    run_flush(run, nb, ngrid, fac, slot, loc);
    A.tall[o] = tall;
    A.norbit[o] = norbit;
    A.steps[o] = steps;
    A.e0[o] = E0;
    A.e1[o] = 0.5 * (vx * vx + vy * vy + vz * vz) + potential64(M, x, y, z);
    A.bbox[o] = bb;
    A.status[o] = status;
    orbit_stats_add(A.stats, steps, status, o);
  }
}

// --------------------------------------------------- instantaneous moments --

// calc_disp_components_for_particle + add_to_energy_grid for every particle
// (main.c:226-229); atomics replace the serial accumulation.
__global__ void k_inst_init(DevModel M, Particles P, double *locI, double *locM, int ngrid) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if(n >= P.n) return;
  const int t = P.type[n];
  double x = P.x[n], y = P.y[n], z = P.z[n];
  double vr2, vt2, vp2, vq2;
  disp_components64(M, t, x, y, z, P.vx[n], P.vy[n], P.vz[n], vr2, vt2, vp2, vq2);
  P.vr2[n] = vr2;
  P.vt2[n] = vt2;
  P.vp2[n] = vp2;
  P.vq2[n] = vq2;
  int iR, iz;
  double fR, fz;
  cell64(x, y, z, M.EG_Rmin, M.EG_logFac, M.EG_Nbin, iR, iz, fR, fz);
  const int nb = M.EG_Nbin;
  const double mass = P.mass[n];
  double w[4] = {(1 - fR) * (1 - fz), (fR) * (1 - fz), (1 - fR) * (fz), (fR) * (fz)};
  int c[4] = {iz * nb + iR, iz * nb + iR + 1, (iz + 1) * nb + iR, (iz + 1) * nb + iR + 1};
  double val[4] = {vr2 * mass, vt2 * mass, vp2 * mass, vq2 * mass};
  double *li = locI + (size_t)t * 4 * ngrid;
  for(int k = 0; k < 4; k++) {
    atomicAdd(&locM[(size_t)t * ngrid + c[k]], w[k] * mass);
    for(int f = 0; f < 4; f++) atomicAdd(&li[(size_t)f * ngrid + c[k]], w[k] * val[f]);
  }
}

// ------------------------------------------------------------------- merit --

// Block-wide sum of up to 8 doubles per thread with a fixed reduction order.
template <int NV>
__device__ void block_sum(double (&v)[NV], double *lds) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for(int k = 0; k < NV; k++) lds[k * nt + tid] = v[k];
  __syncthreads();
  for(int s = nt / 2; s > 0; s >>= 1) {
    if(tid < s)
      for(int k = 0; k < NV; k++) lds[k * nt + tid] += lds[k * nt + tid + s];
    __syncthreads();
  }
  for(int k = 0; k < NV; k++) v[k] = lds[k * nt];
  __syncthreads();
}

// Change of the instantaneous terms (main.c:1125-1176) when the particle's
// squared components change by mass*diff at its energy-grid cell.
__device__ double inst_delta(const DevModel &M, const double *S, const double *W, const int *cutB_of, int t,
                             double x, double y, double z, double mass, const double (&dnew)[4],
                             const double (&dold)[4], double lowdisp, int &c0, double (&cw)[4], int (&cc)[4]) {
  const int L = M.EG_L, nb = M.EG_Nbin, Nstack = GALIC_NSTACK(L);
  int iR, iz;
  double fR, fz;
  cell64(x, y, z, M.EG_Rmin, M.EG_logFac, nb, iR, iz, fR, fz);
  cw[0] = (1 - fR) * (1 - fz);
  cw[1] = (fR) * (1 - fz);
  cw[2] = (1 - fR) * (fz);
  cw[3] = (fR) * (fz);
  cc[0] = iz * nb + iR;
  cc[1] = iz * nb + iR + 1;
  cc[2] = (iz + 1) * nb + iR;
  cc[3] = (iz + 1) * nb + iR + 1;
  c0 = cc[0];
  const double *MR = S + ST_EMR * Nstack, *MT = S + ST_EMT * Nstack;
  const int stI[4] = {ST_IR, ST_IT, ST_IP, ST_IQ};
  const int stE[4] = {ST_ER, ST_ET, ST_EP, ST_EQ};
  int node[4];
  for(int k = 0; k < 4; k++) node[k] = cutB_of[cc[k]];
  double total = 0;
  const int tovs = M.tovs[t];
  for(int f = 0; f < 4; f++) {
    if(f == I_Q && (tovs == 0 || t == 2)) continue;  // main.c:1195
    double val = (dnew[f] - dold[f]) * mass;           // add_to_energy_grid scales by mass first
    const double *I = S + stI[f] * Nstack, *E = S + stE[f] * Nstack;
    for(int k = 0; k < 4; k++) {
      int nd = node[k];
      bool first = true;
      for(int j = 0; j < k; j++)
        if(node[j] == nd) first = false;
      if(!first) continue;
      double delta = 0;
      for(int j = k; j < 4; j++)
        if(node[j] == nd) delta += cw[j] * val;
      double mr = MR[nd], mt = MT[nd];
      if(mr > 0 && mt > 0) {
        double tgt = E[nd] / mt;
        double den = tgt > lowdisp ? tgt : lowdisp;
        double before = fabs(I[nd] / mr - tgt) / den;
        double after = fabs((I[nd] + delta) / mr - tgt) / den;
        total += after - before;
      }
    }
  }
  return total;
}

// One workgroup per trial.  The trial's old and new orbit responses sit in two
// private slots; the change of every orbit-averaged merit term is evaluated on
// the fixed cut of the density target (the cut nodes never change, because the
// targets are fixed): fine-level nodes inside the orbits' bounding box directly,
// coarse nodes from a pyramid of the difference built in LDS.  This equals
// F_try - F_base of main.c:464-465 up to rounding.
__global__ void __launch_bounds__(256) k_eval_apply64(DevModel M, EvalArgs A) {
  extern __shared__ double lds[];
  const int i = A.block_base + blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int L = M.DG_L, nb = M.DG_Nbin, ngrid = nb * nb, Nstack = GALIC_NSTACK(L);
  const int ncoarse = (Nstack - ngrid);
  // coarse levels 0..L-1 in LDS, or in this workgroup's HBM scratch where they do not fit (levels 8, 9)
  double *cs = A.cs_global ? A.cs_global + (size_t)blockIdx.x * A.cs_stride : lds;
  double *red = A.cs_global ? lds : lds + ncoarse;  // reduction scratch, 8*nt
  __shared__ int s_accept;

  const int p = A.tidx[i];
  const int t = A.P.type[p];
  const double *S = A.stacks + (size_t)t * ST_COUNT * Nstack;
  const double *W = A.weights + (size_t)t * W_COUNT * Nstack;
  const int oo = 2 * i - A.slot_base, on = 2 * i + 1 - A.slot_base;
  const double *so = A.slots + (size_t)oo * A.slot_stride;
  const double *sn = A.slots + (size_t)on * A.slot_stride;
  const int4 bo = A.bbox[2 * i], bn = A.bbox[2 * i + 1];
  int r0 = min(bo.x, bn.x), r1 = max(bo.y, bn.y), z0 = min(bo.z, bn.z), z1 = max(bo.w, bn.w);
  if(r1 < r0 || z1 < z0) {  // no binning event (cannot happen after a successful orbit launch)
    r0 = z0 = 0;
    r1 = z1 = -1;
  }
  const int pr0 = r0 >> 1, pr1 = r1 >> 1, pz0 = z0 >> 1, pz1 = z1 >> 1;
  const int npr = (r1 < r0) ? 0 : pr1 - pr0 + 1, npz = (z1 < z0) ? 0 : pz1 - pz0 + 1;
  const int offL = GALIC_STACKOFFSET(L, 0, 0), offL1 = GALIC_STACKOFFSET(L - 1, 0, 0);

  double phi[kFields];
  const int stR[kFields] = {ST_RM, ST_RSR, ST_RST, ST_RSQ, ST_RSP};
  const int stT[kFields] = {ST_MT, ST_ER, ST_ET, ST_EQ, ST_EP};
  (void)npr; (void)npz; (void)pr0; (void)pz0; (void)offL; (void)offL1; (void)pr1; (void)pz1;
  for(int f = 0; f < kFields; f++) {
    CutTerm C;
    C.R = S + stR[f] * Nstack;
    C.T = S + stT[f] * Nstack;
    C.den = W + W_A * Nstack;
    C.ratio = (f == F_M) ? nullptr : W + (W_SR + f - 1) * Nstack;
    const double *fo = so + f, *fn = sn + f;  // cell-major slots: field f of cell c at c*5 + f
    phi[f] = cut_sum_field(C, [&](int cell) { return fn[(size_t)cell * kFields] - fo[(size_t)cell * kFields]; }, r0,
                           r1, z0, z1, L, cs);
  }
  double sums[kFields];
  for(int f = 0; f < kFields; f++) sums[f] = phi[f];
  block_sum<kFields>(sums, red);

  const double x = A.P.x[p], y = A.P.y[p], z = A.P.z[p], mass = A.P.mass[p];
  const double tvx = A.tvx[i], tvy = A.tvy[i], tvz = A.tvz[i];
  double dnew[4], dold[4] = {A.P.vr2[p], A.P.vt2[p], A.P.vp2[p], A.P.vq2[p]};
  if(tid == 0) {
    disp_components64(M, t, x, y, z, tvx, tvy, tvz, dnew[0], dnew[1], dnew[2], dnew[3]);
    const double *fit = A.fit + (size_t)t * 16;
    int c0, cc[4];
    double cw[4];
    double vinst = inst_delta(M, S, W, A.cutB_of + (size_t)t * ngrid, t, x, y, z, mass, dnew, dold, A.lowdisp,
                              c0, cw, cc);
    const double MT = A.MType[t];
    double dF = sums[F_M] / MT +
                (A.fac[t][0] * sums[F_R] + A.fac[t][1] * sums[F_T] + A.fac[t][3] * sums[F_P] +
                 A.fac[t][2] * sums[F_Q]) / MT / 4 +
                fit[5] / fit[6] * vinst;
    int acc = (dF < 0);
    A.accept[i] = acc;
    A.dF[i] = dF;
    s_accept = acc;
    const int ono = A.norbit[2 * i], onn = A.norbit[2 * i + 1];
    A.P.orbits[p] = acc ? onn : ono;
    if(acc) {
      A.P.vx[p] = tvx;
      A.P.vy[p] = tvy;
      A.P.vz[p] = tvz;
      double *lid = A.locI_d + (size_t)t * 4 * ngrid;
      for(int f = 0; f < 4; f++) {
        double val = (dnew[f] - dold[f]) * mass;
        for(int k = 0; k < 4; k++) atomicAdd(&lid[(size_t)f * ngrid + cc[k]], cw[k] * val);
      }
      A.P.vr2[p] = dnew[0];
      A.P.vt2[p] = dnew[1];
      A.P.vp2[p] = dnew[2];
      A.P.vq2[p] = dnew[3];
    }
  }
  __syncthreads();
  const int acc = s_accept;
  // deltas of accepted orbits (main.c:477-486), then clear both slots
  const int nr = (r1 < r0) ? 0 : r1 - r0 + 1, nzz = (z1 < z0) ? 0 : z1 - z0 + 1;
  double *lad = A.locA_d + (size_t)t * kFields * ngrid;
  double *wo = (double *)so, *wn = (double *)sn;
  for(int k = tid; k < nr * nzz; k += nt) {
    int cell = (z0 + k / nr) * nb + (r0 + k % nr);
    for(int f = 0; f < kFields; f++) {
      size_t o = (size_t)cell * kFields + f;
      if(acc) {
        double d = wn[o] - wo[o];
        if(d != 0) atomicAdd(&lad[(size_t)f * ngrid + cell], d);
      }
      wo[o] = 0;
      wn[o] = 0;
    }
  }
}

// ----------------------------------------------------------- commit / fit --

// commit_updates (main.c:808-841): Loc += delta, delta = 0
__global__ void k_commit(double *loc, double *delta, size_t n) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= n) return;
  loc[k] += delta[k];
  delta[k] = 0;
}

// One workgroup per (type, stack): copy the fine level from a Loc grid and
// build the coarse levels as in smooth_stack (grid.c:1247-1265).
__global__ void k_fit_stacks(FitArgs A) {
  const int job = blockIdx.x;  // type*10 + k
  const int t = job / 10 + 1, k = job % 10;
  if(!A.present[t]) return;
  const int L = A.L, nb = 1 << L, ngrid = nb * nb, Nstack = GALIC_NSTACK(L);
  double *S = A.stacks + (size_t)t * ST_COUNT * Nstack;
  const int dst[10] = {ST_RM, ST_RSR, ST_RST, ST_RSQ, ST_RSP, ST_EMR, ST_IR, ST_IT, ST_IP, ST_IQ};
  double *st = S + (size_t)dst[k] * Nstack;
  const double *src;
  if(k < 5)
    src = A.locA + ((size_t)t * kFields + k) * ngrid;
  else if(k == 5)
    src = A.locM + (size_t)t * ngrid;
  else
    src = A.locI + ((size_t)t * 4 + (k - 6)) * ngrid;
  const int off = GALIC_STACKOFFSET(L, 0, 0);
  for(int c = threadIdx.x; c < ngrid; c += blockDim.x) st[off + c] = src[c];
  __syncthreads();
  for(int l = L - 1; l >= 0; l--) {
    int n = 1 << l;
    for(int q = threadIdx.x; q < n * n; q += blockDim.x) {
      int i = q / n, j = q % n;
      st[GALIC_STACKOFFSET(l, i, j)] = st[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j)] +
                                       st[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j)] +
                                       st[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j + 1)] +
                                       st[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j + 1)];
    }
    __syncthreads();
  }
}

// calc_global_fit's stacks for grids of 32 x 32 cells and more, in tiles: workgroup (tile,
// stack) copies the tile's 32 x 32 fine cells into the stack and forms the tile's nodes of
// levels L-1 .. L-5 in LDS, each parent the sum of its four children in smooth_stack's order
// (grid.c:1247-1265); k_fit_top then forms the levels above the tiles.  The same sums in the
// same order as k_fit_stacks, spread over many workgroups.
constexpr int FIT_TILE = 32;
__global__ void __launch_bounds__(256) k_fit_tiles(FitArgs A) {
  __shared__ double lv[2][FIT_TILE * FIT_TILE];
  const int job = blockIdx.y;  // type*10 + k
  const int t = job / 10 + 1, k = job % 10;
  if(!A.present[t]) return;
  const int L = A.L, nb = 1 << L, ngrid = nb * nb, Nstack = GALIC_NSTACK(L), ts = nb / FIT_TILE;
  const int tz = blockIdx.x / ts, tr = blockIdx.x % ts;
  double *S = A.stacks + (size_t)t * ST_COUNT * Nstack;
  const int dst[10] = {ST_RM, ST_RSR, ST_RST, ST_RSQ, ST_RSP, ST_EMR, ST_IR, ST_IT, ST_IP, ST_IQ};
  double *st = S + (size_t)dst[k] * Nstack;
  const double *src;
  if(k < 5)
    src = A.locA + ((size_t)t * kFields + k) * ngrid;
  else if(k == 5)
    src = A.locM + (size_t)t * ngrid;
  else
    src = A.locI + ((size_t)t * 4 + (k - 6)) * ngrid;
  const int off = GALIC_STACKOFFSET(L, 0, 0);
  for(int q = threadIdx.x; q < FIT_TILE * FIT_TILE; q += blockDim.x) {
    const int c = (tz * FIT_TILE + q / FIT_TILE) * nb + tr * FIT_TILE + q % FIT_TILE;
    const double v = src[c];
    st[off + c] = v;
    lv[0][q] = v;
  }
  __syncthreads();
  int cur = 0;
  for(int u = 1; u <= 5; u++) {
    const int n = FIT_TILE >> u, l = L - u;
    const double *ch = lv[cur];
    for(int q = threadIdx.x; q < n * n; q += blockDim.x) {
      const int i = q / n, j = q % n, m = 2 * n;
      const double v = ch[(2 * i) * m + 2 * j] + ch[(2 * i + 1) * m + 2 * j] + ch[(2 * i) * m + 2 * j + 1] +
                       ch[(2 * i + 1) * m + 2 * j + 1];
      lv[1 - cur][q] = v;
      st[GALIC_STACKOFFSET(l, tz * n + i, tr * n + j)] = v;
    }
    cur = 1 - cur;
    __syncthreads();
  }
}

// The levels above the tiles of k_fit_tiles (L-6 .. 0), one workgroup per stack.
__global__ void k_fit_top(FitArgs A) {
  const int job = blockIdx.x;
  const int t = job / 10 + 1, k = job % 10;
  if(!A.present[t]) return;
  const int L = A.L, Nstack = GALIC_NSTACK(L);
  const int dst[10] = {ST_RM, ST_RSR, ST_RST, ST_RSQ, ST_RSP, ST_EMR, ST_IR, ST_IT, ST_IP, ST_IQ};
  double *st = A.stacks + (size_t)t * ST_COUNT * Nstack + (size_t)dst[k] * Nstack;
  for(int l = L - 6; l >= 0; l--) {
    const int n = 1 << l;
    for(int q = threadIdx.x; q < n * n; q += blockDim.x) {
      const int i = q / n, j = q % n;
      st[GALIC_STACKOFFSET(l, i, j)] = st[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j)] +
                                       st[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j)] +
                                       st[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j + 1)] +
                                       st[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j + 1)];
    }
    __syncthreads();
  }
}

// The cut sums of calc_global_fit spread over nblk workgroups per type: partial sums of
// the five merit terms per workgroup, summed in workgroup order by k_fit_final.
__global__ void __launch_bounds__(256) k_fit_partial(FitArgs A, double *part, int nblk) {
  __shared__ double red[5 * 256];
  const int t = blockIdx.y + 1, b = blockIdx.x;
  if(!A.present[t]) return;
  const int L = A.L, Nstack = GALIC_NSTACK(L);
  const double *S = A.stacks + (size_t)t * ST_COUNT * Nstack;
  const double *W = A.weights + (size_t)t * W_COUNT * Nstack;
  const double *den = W + W_A * Nstack, *cutB = W + W_B * Nstack;
  const double *Rm = S + ST_RM * Nstack, *Tm = S + ST_MT * Nstack;
  const double *MR = S + ST_EMR * Nstack, *MT = S + ST_EMT * Nstack;
  const int stI[4] = {ST_IR, ST_IT, ST_IP, ST_IQ};
  const int stE[4] = {ST_ER, ST_ET, ST_EP, ST_EQ};
  double v[5] = {0, 0, 0, 0, 0};
  for(int node = b * blockDim.x + threadIdx.x; node < Nstack; node += nblk * blockDim.x) {
    if(den[node] != 0) v[0] += fabs(Rm[node] - Tm[node]) / den[node];
    if(cutB[node] != 0 && MR[node] > 0 && MT[node] > 0)
      for(int f = 0; f < 4; f++) {
        const double *I = S + stI[f] * Nstack, *E = S + stE[f] * Nstack;
        double tgt = E[node] / MT[node];
        double dd = tgt > A.lowdisp ? tgt : A.lowdisp;
        v[1 + f] += fabs(I[node] / MR[node] - tgt) / dd;
      }
  }
  block_sum<5>(v, red);
  if(threadIdx.x == 0)
    for(int f = 0; f < 5; f++) part[((size_t)blockIdx.y * nblk + b) * 5 + f] = v[f];
}

// The bookkeeping of main.c:294-359 from the partial sums (one thread per type).
__global__ void k_fit_final(FitArgs A, const double *part, int nblk) {
  const int t = blockIdx.x + 1;
  if(!A.present[t] || threadIdx.x != 0) return;
  double v[5] = {0, 0, 0, 0, 0};
  for(int b = 0; b < nblk; b++)
    for(int f = 0; f < 5; f++) v[f] += part[((size_t)blockIdx.x * nblk + b) * 5 + f];
  const int Nstack = GALIC_NSTACK(A.L);
  const double *Tm = A.stacks + (size_t)t * ST_COUNT * Nstack + ST_MT * Nstack;
  double *fit = A.fit + (size_t)t * 16;
  double s = v[0] * (1.0 / Tm[0]);
  double sr = v[1], st = v[2], sp = v[3], sq = v[4];
  int tovs = A.tovs[t];
  double cnt;
  if(tovs == 0 || tovs == 1) {
    cnt = 3.0;
    sq = 0;
  } else
    cnt = 4.0;
  if(fit[8] == 0) {
    fit[5] = s / ((sr + st + sp + sq) / cnt);
    fit[8] = 1;
  }
  double rf = fit[5];
  sr *= rf;
  st *= rf;
  sp *= rf;
  sq *= rf;
  fit[0] = s;
  fit[1] = sr;
  fit[2] = st;
  fit[3] = sp;
  fit[4] = sq;
  fit[6] = cnt;
  fit[7] = s + (sr + st + sp + sq) / cnt;
}

// One workgroup per type: the cut sums of calc_global_fit and the Srelfac
// bookkeeping (main.c:294-359).  fit[t][.] = S, Sdisp_r, Sdisp_t, Sdisp_p,
// Sdisp_q, Srelfac, Srelfac_count, Sall, firstrun flag.
__global__ void k_fit_sums(FitArgs A) {
  extern __shared__ double red[];
  const int t = blockIdx.x + 1;
  if(!A.present[t]) return;
  const int L = A.L, Nstack = GALIC_NSTACK(L);
  const double *S = A.stacks + (size_t)t * ST_COUNT * Nstack;
  const double *W = A.weights + (size_t)t * W_COUNT * Nstack;
  const double *den = W + W_A * Nstack, *cutB = W + W_B * Nstack;
  const double *Rm = S + ST_RM * Nstack, *Tm = S + ST_MT * Nstack;
  const double *MR = S + ST_EMR * Nstack, *MT = S + ST_EMT * Nstack;
  const int stI[4] = {ST_IR, ST_IT, ST_IP, ST_IQ};
  const int stE[4] = {ST_ER, ST_ET, ST_EP, ST_EQ};
  double v[5] = {0, 0, 0, 0, 0};
  for(int node = threadIdx.x; node < Nstack; node += blockDim.x) {
    if(den[node] != 0) v[0] += fabs(Rm[node] - Tm[node]) / den[node];
    if(cutB[node] != 0 && MR[node] > 0 && MT[node] > 0)
      for(int f = 0; f < 4; f++) {
        const double *I = S + stI[f] * Nstack, *E = S + stE[f] * Nstack;
        double tgt = E[node] / MT[node];
        double dd = tgt > A.lowdisp ? tgt : A.lowdisp;
        v[1 + f] += fabs(I[node] / MR[node] - tgt) / dd;
      }
  }
  block_sum<5>(v, red);
  if(threadIdx.x == 0) {
    double *fit = A.fit + (size_t)t * 16;
    double s = v[0] * (1.0 / Tm[0]);
    double sr = v[1], st = v[2], sp = v[3], sq = v[4];
    int tovs = A.tovs[t];
    double cnt;
    if(tovs == 0 || tovs == 1) {
      cnt = 3.0;
      sq = 0;
    } else
      cnt = 4.0;
    if(fit[8] == 0) {
      fit[5] = s / ((sr + st + sp + sq) / cnt);
      fit[8] = 1;
    }
    double rf = fit[5];
    sr *= rf;
    st *= rf;
    sp *= rf;
    sq *= rf;
    fit[0] = s;
    fit[1] = sr;
    fit[2] = st;
    fit[3] = sp;
    fit[4] = sq;
    fit[6] = cnt;
    fit[7] = s + (sr + st + sp + sq) / cnt;
  }
}

// ------------------------------------------------------- trial preparation --

// Orbit o = 2i is the trial's current velocity, o = 2i+1 its trial velocity.
__global__ void k_prepare_trials(int ntrial, const int *tidx, const double *tvx, const double *tvy,
                                 const double *tvz, Particles P, OrbitIn O) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if(o >= 2 * ntrial) return;
  int i = o >> 1, p = tidx[i];
  O.pidx[o] = p;
  O.ox[o] = P.x[p];
  O.oy[o] = P.y[p];
  O.oz[o] = P.z[p];
  O.omass[o] = P.mass[p];
  O.otint[o] = P.tint[p];
  O.otype[o] = P.type[p];
  if(o & 1) {
    O.ovx[o] = tvx[i];
    O.ovy[o] = tvy[i];
    O.ovz[o] = tvz[i];
  } else {
    O.ovx[o] = P.vx[p];
    O.ovy[o] = P.vy[p];
    O.ovz[o] = P.vz[p];
  }
}

// Orbit o = particle (base + o) with its current velocity (initial response pass).
__global__ void k_prepare_all(int n, int base, Particles P, OrbitIn O) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if(o >= n) return;
  int p = base + o;
  O.pidx[o] = p;
  O.ox[o] = P.x[p];
  O.oy[o] = P.y[p];
  O.oz[o] = P.z[p];
  O.ovx[o] = P.vx[p];
  O.ovy[o] = P.vy[p];
  O.ovz[o] = P.vz[p];
  O.omass[o] = P.mass[p];
  O.otint[o] = P.tint[p];
  O.otype[o] = P.type[p];
}

__global__ void k_store_orbits(int n, int base, const int *norbit, int *orbits) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if(o < n) orbits[base + o] = norbit[o];
}

}  // namespace galic
