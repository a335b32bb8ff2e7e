// galic_fast.hip -- the batched FP32 path on gfx950.
//
// Same algorithm as the FP64 path (orbit_response.c:155-491, main.c:365-527,
// 808-841), re-shaped for throughput:
//   * leapfrog and table interpolation in FP32 with the hardware transcendentals
//     (v_log_f32, v_rsq_f32, v_rcp_f32); time and energy bookkeeping in FP64;
//   * the two force tables interleaved as float2 so a bilinear tap is one load;
//   * algebraic forms of the basis vectors: cos(atan2(y,x)) = x/R, and the
//     cross products of orbit_response.c:253-271 written out;
//   * K trial velocities per particle evaluated in one launch, the best taken;
//   * grids accumulated as 64-bit fixed point with integer atomics, so every
//     grid is bit-identical run to run whatever the order of the atomics;
//   * merit sums in FP64 on the same fixed cuts as the FP64 path.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "galic_ctx.h"
#include "galic_fast.h"
#include "galic_kernels.h"
#include "galic_rng.h"

namespace galic {

// ------------------------------------------------------------ FP32 lookups --

__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

struct Fg32 {
  float R, invR, jd, id;
};

__device__ __forceinline__ Fg32 fg_index32(const FastModel &F, float x, float y, float z) {
  Fg32 g;
  float R2 = x * x + y * y;
  g.invR = fast_rsq(R2);
  g.R = R2 * g.invR;
  if(!(R2 > 0)) {
    g.R = 0;
    g.invR = 0;
  }
  g.jd = fast_log2(g.R * F.fg_invRmin + 1.0f) * F.fg_invlog2Fac;
  g.id = fast_log2(fabsf(z) * F.fg_invRmin + 1.0f) * F.fg_invlog2Fac;
  return g;
}

// Every table value one point of the orbit needs, loaded together so a step
// waits on memory once: the four force-table nodes and the four streaming
// velocity nodes around the point (both on the force grid, indices clamped
// into the table; the off-grid force below does not use the force taps), and
// the density-grid cell size that sets the next step's dt (from the LDS copy of
// its per-index table where the grid has one: the size of cell (iR, iz) is the
// larger of the R and z widths, the same function of either index).
struct Taps32 {
  float2 a00, a01, a10, a11;  // (DPotDR, DPotDz) at (z row, R column) +(0,0), (0,1), (1,0), (1,1)
  float v00, v01, v10, v11;   // streaming velocity at the same nodes
  float cell;                 // DG_CellSize of the point's density cell
  int jb, ib;                 // force-grid node left of / below the point (unclamped)
};

__device__ __forceinline__ Taps32 load_taps32(const FastModel &F, const Fg32 &g, int type, int iR, int iz,
                                              const float *s_cell) {
  Taps32 t;
  const int nb = F.fg_nbin;
  t.jb = max((int)g.jd, 0);
  t.ib = max((int)g.id, 0);
  const int k = min(t.ib, nb - 2) * nb + min(t.jb, nb - 2);
  const float2 *T = F.acc + k;
  const float *V = F.vstream + (size_t)type * nb * nb + k;
  t.a00 = T[0];
  t.a01 = T[1];
  t.a10 = T[nb];
  t.a11 = T[nb + 1];
  // (a table of zeros, as the isotropic components have, gives vstr = +0 without its loads)
  if((F.vstream_nz >> type) & 1u) {
    t.v00 = V[0];
    t.v01 = V[1];
    t.v10 = V[nb];
    t.v11 = V[nb + 1];
  } else {
    t.v00 = t.v01 = t.v10 = t.v11 = 0.f;
  }
  t.cell = F.cell1 ? s_cell[max(iR, iz)] : F.cellsize[iz * F.dg_nbin + iR];
  return t;
}

__device__ __forceinline__ void accel32(const FastModel &F, float x, float y, float z, const Fg32 &g,
                                        const Taps32 &t, float &ax, float &ay, float &az) {
  ax = ay = az = 0.f;
  if(F.use_table) {
    const int nb = F.fg_nbin;
    const int jb = t.jb, ib = t.ib;
    if(ib < nb - 1 && jb < nb - 1) {
      float jf = g.jd - jb, fi = g.id - ib;
      float w00 = (1 - fi) * (1 - jf), w01 = (1 - fi) * jf, w10 = fi * (1 - jf), w11 = fi * jf;
      float accR = t.a00.x * w00 + t.a01.x * w01 + t.a10.x * w10 + t.a11.x * w11;
      float accz = t.a00.y * w00 + t.a01.y * w01 + t.a10.y * w10 + t.a11.y * w11;
      if(z < 0) accz = -accz;
      float c = 1.f, s = 0.f;
      if(g.R > 0) {
        c = x * g.invR;
        s = y * g.invR;
      }
      ax = accR * c;
      ay = accR * s;
      az = accz;
    } else {
      float r = sqrtf(x * x + y * y + z * z);
      float f = F.halo_GM * fast_rcp((r + F.halo_eps) * (r + F.halo_a) * (r + F.halo_a)) * F.offgrid_fac;
      ax = -f * x;
      ay = -f * y;
      az = -f * z;
    }
  }
  if(F.add_halo || F.add_bulge) {
    float r = sqrtf(x * x + y * y + z * z);
    float f = 0.f;
    if(F.add_halo) f += F.halo_GM * fast_rcp((r + F.halo_eps) * (r + F.halo_a) * (r + F.halo_a));
    if(F.add_bulge) f += F.bulge_GM * fast_rcp((r + F.bulge_eps) * (r + F.bulge_a) * (r + F.bulge_a));
    ax -= f * x;
    ay -= f * y;
    az -= f * z;
  }
}

// Bilinear streaming velocity; beyond the last node the outermost column / row
// is used (fraction 1 of the clamped cell).
__device__ __forceinline__ float vstream32(const FastModel &F, const Fg32 &g, const Taps32 &t) {
  const int nb = F.fg_nbin;
  float fR = g.jd - t.jb, fz = g.id - t.ib;
  if(t.jb >= nb - 1) fR = 1.f;
  if(t.ib >= nb - 1) fz = 1.f;
  return t.v00 * (1 - fR) * (1 - fz) + t.v10 * (1 - fR) * fz + t.v01 * fR * (1 - fz) + t.v11 * fR * fz;
}

__device__ __forceinline__ void cell32(const FastModel &F, float R, float z, int &iR, int &iz, float &fR,
                                       float &fz) {
  const int nb = F.dg_nbin;
  float br = fast_log2(R * F.dg_invRmin + 1.0f) * F.dg_invlog2Fac - 0.5f;
  float bz = fast_log2(fabsf(z) * F.dg_invRmin + 1.0f) * F.dg_invlog2Fac - 0.5f;
  br = fmaxf(br, 0.f);
  bz = fmaxf(bz, 0.f);
  iR = (int)br;
  iz = (int)bz;
  fR = br - iR;
  fz = bz - iz;
  if(iR >= nb - 1) {
    iR = nb - 2;
    fR = 1.f;
  }
  if(iz >= nb - 1) {
    iz = nb - 2;
    fz = 1.f;
  }
}

// vr2, vt2, vq2, vp2 of orbit_response.c:222-273 in closed form:
// spherical basis Q = (-y, x, 0), T = (-zx, -zy, R^2), |Q| = R, |T| = R r.
__device__ __forceinline__ void decompose32(int tovs, float vstr, float x, float y, float z, float vx, float vy,
                                            float vz, const Fg32 &g, float &vr2, float &vt2, float &vq2, float &vp2) {
  float R2 = x * x + y * y;
  float vq = (vy * x - vx * y) * g.invR;
  vq2 = vq * vq;
  vp2 = (vq - vstr) * (vq - vstr);
  if(tovs == 2) {
    float vdx = vx * x + vy * y;
    vr2 = vdx * vdx * g.invR * g.invR;
    vt2 = vz * vz;
  } else {
    float r2 = R2 + z * z;
    float ir2 = fast_rcp(r2);
    float vdx = vx * x + vy * y + vz * z;
    vr2 = vdx * vdx * ir2;
    float vdt = R2 * vz - z * (vx * x + vy * y);
    vt2 = vdt * vdt * ir2 * g.invR * g.invR;
  }
}

// ------------------------------------------------------------- FP32 orbits --

struct Run32 {
  int quad;
  float w[kFields][4];
};

__device__ __forceinline__ void run32_reset(Run32 &r, int quad) {
  r.quad = quad;
#pragma unroll
  for(int f = 0; f < kFields; f++)
#pragma unroll
    for(int c = 0; c < 4; c++) r.w[f][c] = 0.f;
}

__device__ __forceinline__ void atomic_q(long long *p, double v) {
  atomicAdd((unsigned long long *)p, (unsigned long long)__double2ll_rn(v));
}

// Adds a run's weights to the orbit's slot (cell-major: 5 fields per cell
// contiguous, so a 2x2 block is two short row segments) or, without a slot,
// to the fixed-point Loc grids.
__device__ __forceinline__ void run32_flush(const Run32 &r, int nb, int ngrid, float *slot, long long *locq,
                                            const FastModel &F) {
  if(r.quad < 0) return;
  const int c[4] = {r.quad, r.quad + 1, r.quad + nb, r.quad + nb + 1};
  if(slot) {
    float *s0 = slot + (size_t)c[0] * kFields, *s2 = slot + (size_t)c[2] * kFields;
#pragma unroll
    for(int f = 0; f < kFields; f++) {
      s0[f] += r.w[f][0];
      s0[kFields + f] += r.w[f][1];
      s2[f] += r.w[f][2];
      s2[kFields + f] += r.w[f][3];
    }
  } else {
#pragma unroll
    for(int f = 0; f < kFields; f++) {
      long long *s = locq + (size_t)f * ngrid;
      double sc = F.qscale[f];
#pragma unroll
      for(int k = 0; k < 4; k++) atomic_q(&s[c[k]], sc * (double)r.w[f][k]);
    }
  }
}

// Append-only response log of one orbit: the lane never reads its log back,
// so a flush is a handful of stores and the wave does not wait on memory.
// Arena blocks are claimed LOG_CHUNK at a time per lane and handed to the
// lane's orbits in turn: a claim is an atomic with return on one counter that
// every lane of the chip shares, and its round trip stalls the whole wave.
// Records are written in pairs (indices 2m and 2m + 1, 32 contiguous bytes of
// each field's column of the block): a finished record waits in registers for
// the next one, so the memory system receives whole 32-byte sectors instead of
// half sectors that it would merge with what memory holds.
struct Log32 {
  int n;              // records written
  int blk;            // current arena block
  bool full;          // out of block-table entries or arena blocks
  bool held;          // rec holds a finished record not yet written
  bool qheld;         // q0, q1: the quads of the written pair at n - 2, not yet stored (n % 4 == 2)
  int q0, q1;
  unsigned int next;  // the lane's claimed blocks [next, end)
  unsigned int end;
  float tot[kFields];  // sums of the orbit's record values per field (all values are >= 0)
  Run32 rec;
};

// Claims the arena block that record n starts, if it starts one.
__device__ __forceinline__ void log32_claim(Log32 &lg, int o, const OrbitArgs32 &A) {
  if(lg.full || (lg.n & (LOG_B - 1)) != 0) return;
  const int b = lg.n / LOG_B;
  unsigned int nb = 0xffffffffu;
  if(b < A.log_M) {
    if(lg.next == lg.end) {
      lg.next = atomicAdd(A.log_next, (unsigned int)LOG_CHUNK);
      lg.end = lg.next + LOG_CHUNK;
    }
    nb = lg.next++;
  }
  if(nb >= A.log_cap) {
    lg.full = true;
  } else {
    lg.blk = (int)nb;
    A.log_blocks[(size_t)o * A.log_M + b] = (int)nb;
  }
}

__device__ __forceinline__ void log32_flush(const Run32 &r, Log32 &lg, int o, const OrbitArgs32 &A) {
  if(r.quad < 0) return;
#pragma unroll
  for(int f = 0; f < kFields; f++) lg.tot[f] += (r.w[f][0] + r.w[f][1]) + (r.w[f][2] + r.w[f][3]);
  if(!lg.held) {
    lg.rec = r;
    lg.held = true;
    return;
  }
  log32_claim(lg, o, A);  // n is even: the pair shares a block
  if(!lg.full) {
    // (the quads of two pairs go out as one 16-byte store: a pair at j % 4 == 0 holds its own)
    const int j = lg.n & (LOG_B - 1);
    if((j & 2) == 0) {
      lg.q0 = lg.rec.quad;
      lg.q1 = r.quad;
      lg.qheld = true;
    } else {
      *(int4 *)(A.log_q + (size_t)lg.blk * LOG_B + j - 2) = make_int4(lg.q0, lg.q1, lg.rec.quad, r.quad);
      lg.qheld = false;
    }
    float4 *v = A.log_v + ((size_t)lg.blk * LOG_B + j) * kFields;
#pragma unroll
    for(int f = 0; f < kFields; f++) {
      v[f] = make_float4(lg.rec.w[f][0], lg.rec.w[f][1], lg.rec.w[f][2], lg.rec.w[f][3]);
      v[kFields + f] = make_float4(r.w[f][0], r.w[f][1], r.w[f][2], r.w[f][3]);
    }
  }
  lg.n += 2;
  lg.held = false;
}

// At the orbit's end: the record still held, alone.
__device__ __forceinline__ void log32_finish(Log32 &lg, int o, const OrbitArgs32 &A) {
  // the quads of a pair still held (in the current block: n % 4 == 2)
  if(lg.qheld) {
    if(!lg.full) *(int2 *)(A.log_q + (size_t)lg.blk * LOG_B + ((lg.n - 2) & (LOG_B - 1))) = make_int2(lg.q0, lg.q1);
    lg.qheld = false;
  }
  if(!lg.held) return;
  log32_claim(lg, o, A);
  if(!lg.full) {
    const int j = lg.n & (LOG_B - 1);
    A.log_q[(size_t)lg.blk * LOG_B + j] = lg.rec.quad;
    float4 *v = A.log_v + ((size_t)lg.blk * LOG_B + j) * kFields;
#pragma unroll
    for(int f = 0; f < kFields; f++)
      v[f] = make_float4(lg.rec.w[f][0], lg.rec.w[f][1], lg.rec.w[f][2], lg.rec.w[f][3]);
  }
  lg.n++;
  lg.held = false;
}

// Trapezoid weight h of one point of the orbit (orbit_response.c:283-319),
// already multiplied by mass / timespan (the final scaling of :475-486),
// accumulated in registers while the orbit stays in one 2x2 block.
template <typename Flush>
__device__ __forceinline__ void bin32(Run32 &r, int nb, Flush flush, int iR, int iz, float fR, float fz, float h,
                                      float vr2, float vt2, float vq2, float vp2, int4 &bb) {
  float m[4] = {h * (1 - fR) * (1 - fz), h * fR * (1 - fz), h * (1 - fR) * fz, h * fR * fz};
  int quad = iz * nb + iR;
  if(quad != r.quad) {
    flush(r);
    run32_reset(r, quad);
    bb.x = min(bb.x, iR);
    bb.y = max(bb.y, iR + 1);
    bb.z = min(bb.z, iz);
    bb.w = max(bb.w, iz + 1);
  }
#pragma unroll
  for(int k = 0; k < 4; k++) {
    r.w[F_M][k] += m[k];
    r.w[F_R][k] += m[k] * vr2;
    r.w[F_T][k] += m[k] * vt2;
    r.w[F_Q][k] += m[k] * vq2;
    r.w[F_P][k] += m[k] * vp2;
  }
}

// State of the orbit a lane is integrating.
struct Lane32 {
  int o, type, tovs;
  double timespan, tall;
  float fac;  // mass / timespan, folded into every trapezoid weight
  float *slot;
  long long *locq;
};

// Takes orbits from the launch's counter until one starts cleanly; orbits with a
// non-positive time span are finished at once with GALIC_ERR_ARG, as the FP64
// path does.  Returns false when the launch has no orbits left.
__device__ __forceinline__ bool take_orbit32(const OrbitArgs32 &A, int nb, Lane32 &L, float &x, float &y, float &z,
                                             float &vx, float &vy, float &vz) {
  const int ngrid = nb * nb;
  while(true) {
    unsigned int q = atomicAdd(A.counter, 1u);
    if(q >= (unsigned int)A.n) return false;
    const int o = (int)q;
    L.o = o;
    L.type = A.otype[o];
    L.timespan = A.otint[o];
    L.tall = 0;
    x = (float)A.ox[o];
    y = (float)A.oy[o];
    z = (float)A.oz[o];
    vx = (float)A.ovx[o];
    vy = (float)A.ovy[o];
    vz = (float)A.ovz[o];
    L.fac = (float)(A.omass[o] / L.timespan);
    L.slot = A.slots ? A.slots + (size_t)(o - A.slot_base) * A.slot_stride : nullptr;
    L.locq = A.locq ? A.locq + (size_t)L.type * kFields * ngrid : nullptr;
    if(L.timespan > 0) return true;
    A.tall[o] = 0;
    A.norbit[o] = 0;
    A.steps[o] = 0;
    A.bbox[o] = make_int4(nb, -1, nb, -1);
    A.status[o] = GALIC_ERR_ARG;
    if(A.log_n) {
      A.log_n[o] = 0;
      for(int f = 0; f < kFields; f++) A.log_tot[(size_t)o * kFields + f] = 0.f;
    }
    orbit_stats_add(A.stats, 0, GALIC_ERR_ARG, o);
  }
}

// Persistent orbit integration with per-lane refill: a lane whose orbit ends
// writes it out and takes the next one at once, so the lanes of a wave stay
// busy however uneven the orbit lengths are.  A freshly taken orbit spends one
// pass of the loop on the lookups at its start point (the pre-loop lookups of
// orbit_response.c:199-210); every later pass is one step of :211-458.
template <bool LOG>
__device__ __forceinline__ void orbit32_body(const FastModel &F, const OrbitArgs32 &A, const float *s_cell) {
  const int nb = F.dg_nbin, ngrid = nb * nb;
  Lane32 L;
  Log32 lg;
  lg.n = 0;
  lg.blk = 0;
  lg.full = false;
  lg.held = false;
  lg.qheld = false;
  lg.q0 = lg.q1 = 0;
  lg.next = lg.end = 0;
#pragma unroll
  for(int f = 0; f < kFields; f++) lg.tot[f] = 0.f;
  auto flush = [&](const Run32 &r) {
    if(LOG)
      log32_flush(r, lg, L.o, A);
    else
      run32_flush(r, nb, ngrid, L.slot, L.locq, F);
  };
  float x, y, z, vx, vy, vz;
  if(!take_orbit32(A, nb, L, x, y, z, vx, vy, vz)) return;
  L.tovs = F.tovs[L.type];
  bool fresh = true;
  float ax = 0, ay = 0, az = 0;
  int iR = 0, iz = 0;
  float cellsize = 0.f;  // of the current point's density cell
  float fR = 0, fz = 0, vr2 = 0, vt2 = 0, vq2 = 0, vp2 = 0;
  Run32 run;
  run.quad = -1;
  int4 bb = make_int4(nb, -1, nb, -1);
  float radsign_prev = 0.f;
  int norbit = 0, flag = 0, status = 0;
  long long steps = 0;
  float dt = 0.f, hdt = 0.f;
  // The trapezoid rule weights each point of the orbit by half of each step it
  // ends or starts (orbit_response.c:283-319).  A point is binned once, with
  // both halves: hprev, the half of the step that ended there, is added when
  // the next step's dt is known.
  float hprev = 0.f;
  while(true) {
    if(!fresh) {
      float inv_v = fast_rsq(vx * vx + vy * vy + vz * vz);
      float inv_a = fast_rsq(ax * ax + ay * ay + az * az);
      dt = fminf(F.fo_v200 * inv_a, F.fc * cellsize * inv_v);
      if(L.timespan < L.tall + (double)dt) {
        dt = (float)(L.timespan - L.tall);
        flag = 1;
      }
      if(!(dt >= 0.f && dt < 3.0e38f)) {
        status = GALIC_ERR_ARG;
        if(hprev > 0.f) bin32(run, nb, flush, iR, iz, fR, fz, L.fac * hprev, vr2, vt2, vq2, vp2, bb);
      } else {
        hdt = 0.5f * dt;
        bin32(run, nb, flush, iR, iz, fR, fz, L.fac * (hprev + hdt), vr2, vt2, vq2, vp2, bb);
        vx += hdt * ax;
        vy += hdt * ay;
        vz += hdt * az;
        x += dt * vx;
        y += dt * vy;
        z += dt * vz;
      }
    }
    Fg32 g = fg_index32(F, x, y, z);
    cell32(F, g.R, z, iR, iz, fR, fz);
    const Taps32 taps = load_taps32(F, g, L.type, iR, iz, s_cell);
    float nax, nay, naz;
    accel32(F, x, y, z, g, taps, nax, nay, naz);
    ax = nax;
    ay = nay;
    az = naz;
    if(!fresh && status == 0) {
      vx += hdt * ax;
      vy += hdt * ay;
      vz += hdt * az;
    }
    cellsize = taps.cell;
    decompose32(L.tovs, vstream32(F, g, taps), x, y, z, vx, vy, vz, g, vr2, vt2, vq2, vp2);
    if(fresh) {
      fresh = false;
      continue;
    }
    if(status == 0) {
      hprev = hdt;
      if(flag) bin32(run, nb, flush, iR, iz, fR, fz, L.fac * hprev, vr2, vt2, vq2, vp2, bb);  // the orbit's last point
      L.tall += (double)dt;
      float radsign = vx * x + vy * y + vz * z;
      if(radsign > 0 && radsign_prev < 0) norbit++;
      radsign_prev = radsign;
      steps++;
      if(steps > F.max_steps) status = GALIC_ERR_STEPCAP;
    }
    if(flag || status) {
      flush(run);
      const int o = L.o;
      if(LOG) {
        log32_finish(lg, o, A);
        A.log_n[o] = lg.full ? -1 : lg.n;
        if(lg.full) atomicAdd(&A.stats[3], 1ull);
#pragma unroll
        for(int f = 0; f < kFields; f++) {
          A.log_tot[(size_t)o * kFields + f] = lg.tot[f];
          lg.tot[f] = 0.f;
        }
        lg.n = 0;
        lg.full = false;
      }
      A.tall[o] = L.tall;
      A.norbit[o] = norbit;
      A.steps[o] = steps;
      A.bbox[o] = bb;
      A.status[o] = status;
      if(A.fin) {
        float *fo = A.fin + (size_t)o * 6;
        fo[0] = x;
        fo[1] = y;
        fo[2] = z;
        fo[3] = vx;
        fo[4] = vy;
        fo[5] = vz;
      }
      orbit_stats_add(A.stats, steps, status, o);
      if(!take_orbit32(A, nb, L, x, y, z, vx, vy, vz)) break;
      L.tovs = F.tovs[L.type];
      fresh = true;
      run.quad = -1;
      bb = make_int4(nb, -1, nb, -1);
      radsign_prev = 0.f;
      norbit = 0;
      flag = 0;
      status = 0;
      steps = 0;
      hprev = 0.f;
    }
  }
}

// (the cell-size table into LDS before any lane can leave)
__device__ __forceinline__ void stage_cell1(const FastModel &F, float *s_cell) {
  if(F.cell1)
    for(int i = threadIdx.x; i < F.dg_nbin; i += blockDim.x) s_cell[i] = F.cell1[i];
  __syncthreads();
}

__global__ void __launch_bounds__(256) k_orbit32(FastModel F, OrbitArgs32 A) {
  __shared__ float s_cell[kCell1Max];
  stage_cell1(F, s_cell);
  orbit32_body<false>(F, A, s_cell);
}
__global__ void __launch_bounds__(256) k_orbit32_log(FastModel F, OrbitArgs32 A) {
  __shared__ float s_cell[kCell1Max];
  stage_cell1(F, s_cell);
  orbit32_body<true>(F, A, s_cell);
}

// ------------------------------------------------------------ trial draws --

// K trial velocities per particle.  Trials k < nfresh are fresh draws from the
// distribution of optimize() (main.c:383-435): Gaussians of the target
// dispersions in the local basis, the third around the streaming velocity.
// The other trials perturb the current velocity by sigma_frac times the same
// dispersions.  Both redraw while |v| >= MaxVelInUnitsVesc * Vesc; after 64
// attempts the trial is the current velocity (a no-op trial).
__global__ void k_draw_trials(DevModel M, int ntrial, int K, int nfresh, const int *tidx, Particles P,
                              const double *disp, double sigma_frac, double maxvel, uint32_t seed, uint32_t stream,
                              const int *tovs_of_type, double *vtx, double *vty, double *vtz) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if(q >= ntrial * K) return;
  int i = q / K, k = q % K;
  int p = tidx[i];
  double x = P.x[p], y = P.y[p], z = P.z[p];
  double vx = P.vx[p], vy = P.vy[p], vz = P.vz[p];
  const int type = P.type[p];
  int tovs = tovs_of_type[type];
  double phi = atan2(y, x);
  double sp = sin(phi), cp = cos(phi);
  double e1[3], e2[3], e3[3];
  if(tovs == 2) {
    e1[0] = cp; e1[1] = sp; e1[2] = 0;
    e2[0] = -sp; e2[1] = cp; e2[2] = 0;
    e3[0] = 0; e3[1] = 0; e3[2] = 1;
  } else {
    double theta = acos(z / sqrt(x * x + y * y + z * z));
    double st = sin(theta), ct = cos(theta);
    e1[0] = st * cp; e1[1] = st * sp; e1[2] = ct;
    e2[0] = -ct * cp; e2[1] = -ct * sp; e2[2] = st;
    e3[0] = -sp; e3[1] = cp; e3[2] = 0;
  }
  const bool fresh = k < nfresh;
  const double f = fresh ? 1.0 : sigma_frac;
  double s1 = f * sqrt(disp[4 * (size_t)p + 0]);
  double s2 = f * sqrt(disp[4 * (size_t)p + 1]);
  double s3 = f * sqrt(disp[4 * (size_t)p + 2]);
  // fresh draws are centred on (0, 0, v_str) in the local basis, perturbations on v
  double cx = vx, cy = vy, cz = vz;
  if(fresh) {
    double vstr = vstream64(M, fg_index64(M, x, y, z), type);
    cx = vstr * e3[0];
    cy = vstr * e3[1];
    cz = vstr * e3[2];
  }
  double lim = maxvel * P.vesc[p];
  double ox = vx, oy = vy, oz = vz;
  for(uint32_t attempt = 0; attempt < 64; attempt++) {
    double g[4];
    gauss4((uint64_t)p, attempt, (uint32_t)k, seed, stream, g);
    double a = g[0] * s1, b = g[1] * s2, c = g[2] * s3;
    double tx = cx + a * e1[0] + b * e2[0] + c * e3[0];
    double ty = cy + a * e1[1] + b * e2[1] + c * e3[1];
    double tz = cz + a * e1[2] + b * e2[2] + c * e3[2];
    if(sqrt(tx * tx + ty * ty + tz * tz) < lim) {
      ox = tx;
      oy = ty;
      oz = tz;
      break;
    }
  }
  vtx[q] = ox;
  vty[q] = oy;
  vtz[q] = oz;
}

// Orbits of a K-trial chunk: orbit j*(K+1) is the particle of trial i = order[j]
// at its current velocity, orbits j*(K+1)+1+k its trials (i*K+k).  Without an
// order the chunk is trials 0..ntrial-1 in turn.
__global__ void k_prepare_ktrials(int ntrial, int K, const int *order, const int *tidx, const double *vtx,
                                  const double *vty, const double *vtz, Particles P, OrbitIn O) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if(o >= ntrial * (K + 1)) return;
  int j = o / (K + 1), k = o % (K + 1);
  int i = order ? order[j] : j;
  int p = tidx[i];
  O.pidx[o] = p;
  O.ox[o] = P.x[p];
  O.oy[o] = P.y[p];
  O.oz[o] = P.z[p];
  O.omass[o] = P.mass[p];
  O.otint[o] = P.tint[p];
  O.otype[o] = P.type[p];
  if(k == 0) {  // (also the whole K = 0 case, where the trial arrays are unused)
    O.ovx[o] = P.vx[p];
    O.ovy[o] = P.vy[p];
    O.ovz[o] = P.vz[p];
  } else {
    size_t q = (size_t)i * K + (k - 1);
    O.ovx[o] = vtx[q];
    O.ovy[o] = vty[q];
    O.ovz[o] = vtz[q];
  }
}

// Energy at the start and at the end of each FP32 orbit, with the FP64
// potential (orbit_response.c:201-208, 464-473).
__global__ void k_energy32(DevModel M, int n, const double *ox, const double *oy, const double *oz, const double *ovx,
                           const double *ovy, const double *ovz, const float *fin, double *e0, double *e1) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if(o >= n) return;
  e0[o] = 0.5 * (ovx[o] * ovx[o] + ovy[o] * ovy[o] + ovz[o] * ovz[o]) + potential64(M, ox[o], oy[o], oz[o]);
  const float *f = fin + (size_t)o * 6;
  const double x = f[0], y = f[1], z = f[2], vx = f[3], vy = f[4], vz = f[5];
  e1[o] = 0.5 * (vx * vx + vy * vy + vz * vz) + potential64(M, x, y, z);
}

// inv[ord[j]] = j: chunk position of each trial.
__global__ void k_invert_order(int n, const int *ord, int *inv) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if(j < n) inv[ord[j]] = j;
}

// Expected orbit length of each trial's particle: the steps its current orbit took.
__global__ void k_step_keys(int ntrial, const int *tidx, const int *psteps, unsigned int *key, int *ord) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= ntrial) return;
  key[i] = (unsigned int)psteps[tidx[i]];
  ord[i] = i;
}

__global__ void k_store_steps(int n, const long long *osteps, int *psteps) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if(o < n) psteps[o] = (int)min(osteps[o], (long long)0x7fffffff);
}

// Records of each trial's K+1 logs (the merit kernel's work for it).
__global__ void k_record_keys(int n, int K, const int *slot_pos, const int *log_n, unsigned int *key, int *ord) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if(b >= n) return;
  const size_t o0 = (size_t)(slot_pos ? slot_pos[b] : b) * (K + 1);
  unsigned int s = 0;
  for(int o = 0; o <= K; o++) s += (unsigned int)max(log_n[o0 + o], 0);
  key[b] = s;
  ord[b] = b;
}

int sort_desc(int n, unsigned int *key, unsigned int *key2, int *ord, int *ord2, void **tmp, size_t *tmp_bytes,
              hipStream_t stream);

__global__ void k_fill_u32(unsigned int *p, unsigned int v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = v;
}

// The order of k_record_keys + sort_desc in one workgroup: (inverted key, index) pairs in LDS
// (dynamic, 8 bytes per element of the power of two above n), bitonic sort ascending.
__global__ void __launch_bounds__(1024) k_record_order(int n, int K, const int *slot_pos, const int *log_n, int *ord) {
  extern __shared__ unsigned long long ko[];
  int m = 1;
  while(m < n) m <<= 1;
  for(int b = threadIdx.x; b < m; b += blockDim.x) {
    unsigned long long v = ~0ull;
    if(b < n) {
      const size_t o0 = (size_t)slot_pos[b] * (K + 1);
      unsigned int s = 0;
      for(int o = 0; o <= K; o++) s += (unsigned int)max(log_n[o0 + o], 0);
      v = ((unsigned long long)(~s) << 32) | (unsigned int)b;
    }
    ko[b] = v;
  }
  __syncthreads();
  for(int size = 2; size <= m; size <<= 1)
    for(int stride = size >> 1; stride > 0; stride >>= 1) {
      for(int i = threadIdx.x; i < m / 2; i += blockDim.x) {
        const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const unsigned long long a = ko[lo], c = ko[hi];
        if((a > c) == up) {
          ko[lo] = c;
          ko[hi] = a;
        }
      }
      __syncthreads();
    }
  for(int b = threadIdx.x; b < n; b += blockDim.x) ord[b] = (int)(unsigned int)ko[b];
}

// Trials in order of decreasing expected orbit length (stable, so the order is
// reproducible): chunks then hold orbits of similar length and the longest
// start first.
int sort_trials_by_steps(int ntrial, const int *tidx, const int *psteps, unsigned int *key, unsigned int *key2,
                         int *ord, int *ord2, void **tmp, size_t *tmp_bytes, hipStream_t stream) {
  if(ntrial <= 0) return 0;
  hipLaunchKernelGGL(k_step_keys, dim3((ntrial + 255) / 256), dim3(256), 0, stream, ntrial, tidx, psteps, key, ord);
  if(hipGetLastError() != hipSuccess) return -1;
  return sort_desc(ntrial, key, key2, ord, ord2, tmp, tmp_bytes, stream);
}

// The trials of a merit launch in order of decreasing record count (ord2): workgroups that take
// them in this order from a shared counter finish together (longest first).
int sort_trials_by_records(int n, int K, const int *slot_pos, const int *log_n, unsigned int *key, unsigned int *key2,
                           int *ord, int *ord2, void **tmp, size_t *tmp_bytes, hipStream_t stream) {
  if(n <= 0) return 0;
  hipLaunchKernelGGL(k_record_keys, dim3((n + 255) / 256), dim3(256), 0, stream, n, K, slot_pos, log_n, key, ord);
  if(hipGetLastError() != hipSuccess) return -1;
  return sort_desc(n, key, key2, ord, ord2, tmp, tmp_bytes, stream);
}

// (key, ord) pairs sorted by decreasing key into (key2, ord2); stable.
int sort_desc(int ntrial, unsigned int *key, unsigned int *key2, int *ord, int *ord2, void **tmp, size_t *tmp_bytes,
              hipStream_t stream) {
  size_t need = 0;
  if(rocprim::radix_sort_pairs_desc(nullptr, need, key, key2, ord, ord2, (unsigned int)ntrial, 0, 32, stream) !=
     hipSuccess)
    return -1;
  if(need > *tmp_bytes) {
    if(*tmp) (void)hipFree(*tmp);
    *tmp = nullptr;
    *tmp_bytes = 0;
    if(hipMalloc(tmp, need) != hipSuccess) return -1;
    *tmp_bytes = need;
  }
  if(rocprim::radix_sort_pairs_desc(*tmp, need, key, key2, ord, ord2, (unsigned int)ntrial, 0, 32, stream) !=
     hipSuccess)
    return -1;
  return 0;
}

// -------------------------------------------------------------- FP32 merit --

// Change of the instantaneous merit terms, as inst_delta of the FP64 path.
__device__ double inst_delta_q(const DevModel &M, const double *S, const int *cutB_of, int t, double x, double y,
                               double z, double mass, const double (&dnew)[4], const double (&dold)[4],
                               double lowdisp, double (&cw)[4], int (&cc)[4]) {
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
  const double *MR = S + ST_EMR * Nstack, *MT = S + ST_EMT * Nstack;
  const int stI[4] = {ST_IR, ST_IT, ST_IP, ST_IQ};
  const int stE[4] = {ST_ER, ST_ET, ST_EP, ST_EQ};
  int node[4];
  for(int k = 0; k < 4; k++) node[k] = cutB_of[cc[k]];
  double total = 0;
  const int tovs = M.tovs[t];
  for(int f = 0; f < 4; f++) {
    if(f == I_Q && (tovs == 0 || t == 2)) continue;
    double val = (dnew[f] - dold[f]) * mass;
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
        total += (fabs((I[nd] + delta) / mr - tgt) - fabs(I[nd] / mr - tgt)) / den;
      }
    }
  }
  return total;
}

// Node terms of the orbit-averaged merit, refreshed after every commit: b =
// response - target and w = the node's weight, 1/min(dist, R200/20) for the
// mass term and (Tm/E_x)/min(dist, R200/20) for an S term on the cut nodes, 0
// elsewhere (calc_stack_difference_mod flags 0 and 2, grid.c:1115-1151).
__global__ void k_pack_merit(const double *stacks, const double *weights, double2 *bw, float2 *bwf, int Nstack) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if(q >= GALIC_NTYPES * kFields * Nstack) return;
  const int node = q % Nstack, f = (q / Nstack) % kFields, t = q / (Nstack * kFields);
  const int stR[kFields] = {ST_RM, ST_RSR, ST_RST, ST_RSQ, ST_RSP};
  const int stT[kFields] = {ST_MT, ST_ER, ST_ET, ST_EQ, ST_EP};
  const double *S = stacks + (size_t)t * ST_COUNT * Nstack;
  const double *W = weights + (size_t)t * W_COUNT * Nstack;
  const double den = W[W_A * Nstack + node];
  const double ratio = (f == F_M) ? 1.0 : W[(W_SR + f - 1) * Nstack + node];
  double2 v;
  v.x = S[stR[f] * Nstack + node] - S[stT[f] * Nstack + node];
  v.y = (den != 0 && ratio != 0) ? ratio / den : 0.0;
  bw[q] = v;
  if(bwf) bwf[q] = make_float2((float)v.x, (float)v.y);
}

__device__ __forceinline__ double merit_term(const double2 *bw, int node, double d) {
  const double2 v = bw[node];
  return (v.y != 0 && d != 0) ? (fabs(v.x + d) - fabs(v.x)) * v.y : 0.0;
}

// Records of one orbit's log held by a thread: r = tid + u * blockDim.x.
constexpr int LOG_RPT = 2;
struct LogRegs {
  int quad[LOG_RPT];
  float4 x[LOG_RPT];
};

// Loads field f of the first LOG_RPT * blockDim.x records of a log (tbl: the
// orbit's arena blocks) into registers.
__device__ __forceinline__ void log_load(const EvalArgs32 &A, const int *tbl, int n, int f, LogRegs &R) {
#pragma unroll
  for(int u = 0; u < LOG_RPT; u++) {
    const int r = threadIdx.x + u * blockDim.x;
    R.quad[u] = -1;
    if(r < n) {
      const int blk = tbl[r / LOG_B], jj = r & (LOG_B - 1);
      R.quad[u] = A.log_q[(size_t)blk * LOG_B + jj];
      R.x[u] = A.log_v[((size_t)blk * LOG_B + jj) * kFields + f];
    }
  }
}

// Each FP32 value times a power of two is rounded to an int64 and added with an
// integer atomic, so the sums are exact and independent of the order.
// quad = iz * nb + iR names the 2x2 cells (iz..iz+1, iR..iR+1), nb = 2^L; the
// LDS grid holds cell (iz, iR) at iz * rs + iR.
// Marks the level L-1 nodes (2 x 2 cells) that the 2x2 cells of quad fall in
// (grids above 128^2, where the pyramid and the delta pass read only the cells
// of marked nodes); nb / 2 nodes per row.
__device__ __forceinline__ void mark_nodes(unsigned int *mask, int quad, int nb, int L) {
  const int iz = quad >> L, iR = quad & (nb - 1), nbp = nb >> 1;
  const int z0 = iz >> 1, z1 = (iz + 1) >> 1, r0 = iR >> 1, r1 = (iR + 1) >> 1;
  const int b[4] = {z0 * nbp + r0, z0 * nbp + r1, z1 * nbp + r0, z1 * nbp + r1};
#pragma unroll
  for(int k = 0; k < 4; k++) {
    const unsigned int bit = 1u << (b[k] & 31);
    if(!(mask[b[k] >> 5] & bit)) atomicOr(&mask[b[k] >> 5], bit);
  }
}

__device__ __forceinline__ void add_record(unsigned long long *G, int quad, float4 x, float sc, int nb, int L, int rs) {
  unsigned long long *g = G + (quad >> L) * rs + (quad & (nb - 1));
  atomicAdd(g, (unsigned long long)__float2ll_rn(x.x * sc));
  atomicAdd(g + 1, (unsigned long long)__float2ll_rn(x.y * sc));
  atomicAdd(g + rs, (unsigned long long)__float2ll_rn(x.z * sc));
  atomicAdd(g + rs + 1, (unsigned long long)__float2ll_rn(x.w * sc));
}

// Adds sign * field f of a whole log into G: the records held in R, then the rest.
__device__ __forceinline__ void log_add(const EvalArgs32 &A, const int *tbl, int n, int f, const LogRegs &R, int nb,
                                        int L, int rs, float sign, unsigned long long *G, unsigned int *mask = nullptr) {
  const float sc = sign * A.log_scale[f];
#pragma unroll
  for(int u = 0; u < LOG_RPT; u++)
    if(R.quad[u] >= 0) {
      add_record(G, R.quad[u], R.x[u], sc, nb, L, rs);
      if(mask) mark_nodes(mask, R.quad[u], nb, L);
    }
  for(int r = threadIdx.x + LOG_RPT * blockDim.x; r < n; r += blockDim.x) {
    const int blk = tbl[r / LOG_B], jj = r & (LOG_B - 1);
    const int quad = A.log_q[(size_t)blk * LOG_B + jj];
    add_record(G, quad, A.log_v[((size_t)blk * LOG_B + jj) * kFields + f], sc, nb, L, rs);
    if(mask) mark_nodes(mask, quad, nb, L);
  }
}

// Zeroes the four cells of every level L-1 node marked in mask (row stride nb;
// nmw mask words, nb / 2 nodes per row).
__device__ __forceinline__ void clear_marked(long long *G, int nb, const unsigned int *mask, int nmw) {
  const int nbp = nb >> 1;
  for(int q = threadIdx.x; q < nmw; q += blockDim.x)
    for(unsigned int bits = mask[q]; bits; bits &= bits - 1) {
      const int nd = q * 32 + __builtin_ctz(bits), iz = 2 * (nd / nbp), iR = 2 * (nd % nbp);
      long long *g = G + iz * nb + iR;
      g[0] = 0;
      g[1] = 0;
      g[nb] = 0;
      g[nb + 1] = 0;
    }
}

// Zeroes G (row stride rs) over the box of fine cells.
__device__ __forceinline__ void clear_box(long long *G, int rs, int4 b) {
  const int nr = (b.y < b.x) ? 0 : b.y - b.x + 1, nzz = (b.w < b.z) ? 0 : b.w - b.z + 1;
  for(int q = threadIdx.x; q < nr * nzz; q += blockDim.x) G[(b.z + q / nr) * rs + b.x + q % nr] = 0;
}

__device__ __forceinline__ int morton_x(int lane) { return (lane & 1) | ((lane >> 1) & 2) | ((lane >> 2) & 4); }
__device__ __forceinline__ int morton_y(int lane) { return ((lane >> 1) & 1) | ((lane >> 2) & 2) | ((lane >> 3) & 4); }

// Sums over groups of 4, 16 and 64 lanes, with lanes in Morton order the
// children of one node.  Within a row of 16 lanes the data moves with DPP
// (quad permutations, row rotations), across rows with readlane: no LDS
// permute, whose round trip would sit on the pyramid's dependency chain.
constexpr int kDppQuadXor1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int kDppRowRor4 = 0x124;
constexpr int kDppRowRor8 = 0x128;

template <int CTRL>
__device__ __forceinline__ long long dpp_i64(long long v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned int)v, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), CTRL, 0xF, 0xF, false);
  return (long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  return __longlong_as_double(dpp_i64<CTRL>(__double_as_longlong(v)));
}
__device__ __forceinline__ long long readlane_i64(long long v, int lane) {
  const int lo = __builtin_amdgcn_readlane((int)(unsigned int)v, lane);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
  return (long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo);
}
// step 0: every lane gets its quad's sum; step 1: its row's (16 lanes); step 2:
// the wave's (64 lanes).  The groups of step s are the sums of step s - 1.
// All 64 lanes of the wave must be active.
__device__ __forceinline__ long long group_sum_i64(long long s, int step) {
  if(step == 0) {
    s += dpp_i64<kDppQuadXor1>(s);
    s += dpp_i64<kDppQuadXor2>(s);
  } else if(step == 1) {
    s += dpp_i64<kDppRowRor4>(s);
    s += dpp_i64<kDppRowRor8>(s);
  } else {
    s = readlane_i64(s, 0) + readlane_i64(s, 16) + readlane_i64(s, 32) + readlane_i64(s, 48);
  }
  return s;
}

__device__ __forceinline__ double term_at(const double2 &v, double d) {
  return (v.y != 0 && d != 0) ? (fabs(v.x + d) - fabs(v.x)) * v.y : 0.0;
}

// The part of the merit pyramid one lane evaluates.  The grid is cut into
// blocks of 16 x 16 fine cells (the whole grid below L = 4); a wave takes blocks
// wave, wave + nwaves, ...  In a block each lane owns a 2x2 cell node, lanes in
// Morton order, so that groups of 4, 16 and 64 lanes are the node's parent,
// grandparent and the block node, summed with group_sum_i64.  The levels above
// the blocks are formed from the block sums afterwards (merit_top up to 8 x 8
// blocks, merit_top_levels above).  The (residual, weight) pairs of fine cells
// and level L-1 nodes come from memory (bw), those of the coarser levels from
// bwc (LDS, indexed by node, up to 128^2 cells; memory above).
__device__ __forceinline__ void grid_blocks(int L, int &bside, int &nblk, int &lblk) {
  const int LB = L < 4 ? L : 4;
  bside = 1 << LB;
  nblk = (1 << L) / bside;
  lblk = L - LB;
}

// Block phase: this lane's terms of D = G - O (D = G without O; G cleared as it
// is read), the block sums to topb.  Returns the lane's partial (without coef).
__device__ __forceinline__ double merit_block(const double2 *bw, const double2 *bwc, int L, int rs, double unit,
                                              long long *G, const long long *O, int4 box, long long *topb,
                                              const unsigned int *maskO = nullptr,
                                              const unsigned int *maskG = nullptr) {
  const int nb = 1 << L, lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int bside, nblk, lblk;
  grid_blocks(L, bside, nblk, lblk);
  const int offL = GALIC_STACKOFFSET(L, 0, 0);
  const int jx = morton_x(lane), jy = morton_y(lane);
  const bool valid = 2 * jy < bside && 2 * jx < bside;
  double acc = 0;
  for(int blk = wave; blk < nblk * nblk; blk += nw) {
    const int bz = blk / nblk, br = blk % nblk;
    bool hit = box.y >= box.x && box.w >= box.z &&
               !(box.y < br * bside || box.x > br * bside + bside - 1 || box.w < bz * bside ||
                 box.z > bz * bside + bside - 1);
    long long s = 0;
    const int pz = bz * (bside >> 1) + jy, pr = br * (bside >> 1) + jx;
    // with masks: only the lanes whose node either log touched read memory (the
    // other nodes' cells are zero in both grids), and a block without one is skipped
    bool mine = valid;
    if(maskG) {
      const int nd = pz * (nb >> 1) + pr;
      mine = valid && ((((maskO[nd >> 5] | maskG[nd >> 5]) >> (nd & 31)) & 1u) != 0);
      hit = hit && __ballot(mine) != 0;
    }
    if(hit) {
      // with masks the grids and all pairs are in memory: the pairs of the block's
      // coarse levels (levels L-2 .. L-4, used by the group leaders) are loaded with
      // the fine ones, so a block costs one round trip to memory
      double2 wc[3];
      if(maskG)
#pragma unroll
        for(int u = 0; u < 3; u++) {
          const int l = L - 2 - u, sh = L - 1 - l;
          if(l >= lblk) wc[u] = bwc[GALIC_STACKOFFSET(l, pz >> sh, pr >> sh)];
        }
      if(mine) {
        int cell[4];
        double2 w[5];
#pragma unroll
        for(int c = 0; c < 4; c++) {
          cell[c] = (2 * pz + (c & 1)) * rs + 2 * pr + (c >> 1);
          w[c] = bw[offL + (2 * pz + (c & 1)) * nb + 2 * pr + (c >> 1)];
        }
        w[4] = bw[GALIC_STACKOFFSET(L - 1, pz, pr)];
#pragma unroll
        for(int c = 0; c < 4; c++) {
          const long long d = O ? G[cell[c]] - O[cell[c]] : G[cell[c]];
          G[cell[c]] = 0;
          s += d;
          acc += term_at(w[c], (double)d * unit);
        }
        acc += term_at(w[4], (double)s * unit);
      }
#pragma unroll
      for(int u = 0; u < 3; u++) {
        const int l = L - 2 - u, g = 2 * u + 1;
        if(l < lblk) break;
        s = group_sum_i64(s, u);
        const int mask = (1 << (g + 1)) - 1, sh = L - 1 - l;
        if((lane & mask) == 0 && valid)
          acc += term_at(maskG ? wc[u] : bwc[GALIC_STACKOFFSET(l, pz >> sh, pr >> sh)], (double)s * unit);
      }
    }
    if(lane == 0) topb[blk] = s;  // 0 for blocks the box misses
  }
  return acc;
}

// Levels above the blocks (wave 0; at most 8 x 8 blocks).
__device__ __forceinline__ double merit_top(const double2 *bwc, int L, double unit, const long long *topb) {
  const int lane = threadIdx.x & 63;
  int bside, nblk, lblk;
  grid_blocks(L, bside, nblk, lblk);
  if(lblk == 0 || (threadIdx.x >> 6) != 0) return 0.0;
  const int bx = morton_x(lane), by = morton_y(lane);
  const bool in = bx < nblk && by < nblk;
  long long s = in ? topb[by * nblk + bx] : 0;
  double acc = 0;
#pragma unroll
  for(int u = 0; u < 3; u++) {
    const int l = lblk - 1 - u, g = 2 * u + 1;
    if(l < 0) break;
    s = group_sum_i64(s, u);
    const int mask = (1 << (g + 1)) - 1, sh = lblk - l;
    if((lane & mask) == 0 && in) acc += term_at(bwc[GALIC_STACKOFFSET(l, by >> sh, bx >> sh)], (double)s * unit);
  }
  return acc;
}

// Levels above the blocks for grids of more than 8 x 8 blocks (above 128^2
// cells): level by level through the pyramid of block sums held in LDS (top,
// stack layout of levels 0..lblk, block sums at level lblk); every thread takes
// part and the function ends at a barrier.
__device__ double merit_top_levels(const double2 *bw, int lblk, double unit, long long *top) {
  double acc = 0;
  for(int l = lblk - 1; l >= 0; l--) {
    const int n = 1 << l;
    for(int q = threadIdx.x; q < n * n; q += blockDim.x) {
      const int i = q / n, j = q % n;
      const long long s = top[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j)] + top[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j)] +
                          top[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j + 1)] +
                          top[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j + 1)];
      top[GALIC_STACKOFFSET(l, i, j)] = s;
      acc += term_at(bw[GALIC_STACKOFFSET(l, i, j)], (double)s * unit);
    }
    __syncthreads();
  }
  return acc;
}

// The sum over the wave's 64 lanes (every lane active), in every lane.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<kDppQuadXor1>(v);
  v += dpp_f64<kDppQuadXor2>(v);
  v += dpp_f64<kDppRowRor4>(v);
  v += dpp_f64<kDppRowRor8>(v);
  const long long b = __double_as_longlong(v);
  return (__longlong_as_double(readlane_i64(b, 0)) + __longlong_as_double(readlane_i64(b, 16))) +
         (__longlong_as_double(readlane_i64(b, 32)) + __longlong_as_double(readlane_i64(b, 48)));
}

// The instantaneous terms of each trial (one trial per thread, inst_delta_q), the total
// merit change F_try - F_base per trial (part: the orbit-averaged terms per trial and wave)
// and the choice: the trial of the smallest negative change is taken, with its
// instantaneous deltas and the particle's new velocity and moments (main.c:464-527 for the
// best of K).  Returns the chosen trial (-1: none) in every thread; ends at a barrier.
__device__ int choose_trial(const DevModel &M, const EvalArgs32 &A, int ti, int p, int t, size_t o0,
                            const double *part, double *s_dF, int *s_best) {
  const int tid = threadIdx.x, nt = blockDim.x, K = A.K;
  const int nb = M.EG_Nbin, ngrid = nb * nb, Nstack = GALIC_NSTACK(M.DG_L);
  const double *S = A.stacks + (size_t)t * ST_COUNT * Nstack;
  const double x = A.P.x[p], y = A.P.y[p], z = A.P.z[p], mass = A.P.mass[p];
  const double dold[4] = {A.P.vr2[p], A.P.vt2[p], A.P.vp2[p], A.P.vq2[p]};
  for(int k = tid; k < K; k += nt) {
    const size_t q = (size_t)ti * K + k;
    double dnew[4];
    disp_components64(M, t, x, y, z, A.tvx[q], A.tvy[q], A.tvz[q], dnew[0], dnew[1], dnew[2], dnew[3]);
    double cw[4];
    int cc[4];
    double vinst = inst_delta_q(M, S, A.cutB_of + (size_t)t * ngrid, t, x, y, z, mass, dnew, dold, A.lowdisp, cw, cc);
    const double *fit = A.fit + (size_t)t * 16;
    double orbit_terms = 0;
    for(int w = 0; w < MERIT_MAX_WAVES; w++) orbit_terms += part[k * MERIT_MAX_WAVES + w];
    s_dF[k] = orbit_terms + fit[5] / fit[6] * vinst;
    A.dFk[q] = s_dF[k];
  }
  __syncthreads();
  if(tid == 0) {
    int best = -1;
    double bv = 0;
    for(int k = 0; k < K; k++)
      if(s_dF[k] < bv) {
        bv = s_dF[k];
        best = k;
      }
    *s_best = best;
    A.accept[ti] = best + 1;
    A.dF[ti] = (best >= 0) ? bv : (K > 0 ? s_dF[0] : 0.0);
    const size_t oc = (best >= 0) ? o0 + 1 + best : o0;
    A.P.orbits[p] = A.norbit[oc];
    A.psteps[p] = (int)min(A.osteps[oc], (long long)0x7fffffff);
    if(best >= 0) {
      const size_t q = (size_t)ti * K + best;
      const double vt[3] = {A.tvx[q], A.tvy[q], A.tvz[q]};
      double dnew[4];
      disp_components64(M, t, x, y, z, vt[0], vt[1], vt[2], dnew[0], dnew[1], dnew[2], dnew[3]);
      int iR, iz;
      double fR, fz;
      cell64(x, y, z, M.EG_Rmin, M.EG_logFac, M.EG_Nbin, iR, iz, fR, fz);
      double cw[4] = {(1 - fR) * (1 - fz), fR * (1 - fz), (1 - fR) * fz, fR * fz};
      int cc[4] = {iz * nb + iR, iz * nb + iR + 1, (iz + 1) * nb + iR, (iz + 1) * nb + iR + 1};
      long long *lid = A.locI_dq + (size_t)t * 4 * ngrid;
      for(int f = 0; f < 4; f++) {
        double val = (dnew[f] - dold[f]) * mass;
        for(int c = 0; c < 4; c++) atomic_q(&lid[(size_t)f * ngrid + cc[c]], A.iscale * cw[c] * val);
      }
      A.P.vx[p] = vt[0];
      A.P.vy[p] = vt[1];
      A.P.vz[p] = vt[2];
      A.P.vr2[p] = dnew[0];
      A.P.vt2[p] = dnew[1];
      A.P.vp2[p] = dnew[2];
      A.P.vq2[p] = dnew[3];
    }
  }
  __syncthreads();
  return *s_best;
}

// One particle per workgroup: the merit change of each of its K trials (the
// orbit-averaged terms on the fixed density cut, the instantaneous terms on the
// fixed cut of EGs_MassResponse), the best improving trial taken and its deltas
// added in fixed point.  Field by field, the current orbit's log is summed into
// O once and each trial's into G, so D = G - O exactly (USE_O); where two grids
// do not fit in LDS (128^2 cells) both logs go into G for every trial.
// LDS: part [K][8] double | topb [64] int64 | bwc [coarse nodes] double2 | G [nb rows] int64 |
//      O [nb rows] int64 (USE_O) | tbl [K+1][log_M] int; a row of G and O holds rs cells
template <bool USE_O, bool GMEM, int LF>
__global__ void __launch_bounds__(merit_threads(LF)) k_eval_log32(DevModel M, EvalArgs32 A) {
  extern __shared__ double lds_d[];
  const int tid = threadIdx.x, nt = blockDim.x, K = A.K, wave = tid >> 6, lane = tid & 63, nw = nt >> 6;
  // LF: the grid level fixed at compile time (0: M.DG_L), so that the index arithmetic of
  // the pyramid folds
  const int L = LF ? LF : M.DG_L, nb = 1 << L, ngrid = nb * nb, Nstack = GALIC_NSTACK(L);
  const int ncb = GALIC_NSTACK(L - 2);  // nodes of levels 0..L-2
  int bside, nblk, lblk;
  grid_blocks(L, bside, nblk, lblk);
  double *part = lds_d;
  long long *top = (long long *)(part + MERIT_MAX_WAVES * K);  // GMEM: the pyramid of block sums, levels 0..lblk
  long long *topb = GMEM ? top + GALIC_STACKOFFSET(lblk, 0, 0) : top;
  const int ntop = GMEM ? GALIC_NSTACK(lblk) : 64;
  double2 *bwc_lds = (double2 *)(top + ntop);
  // GMEM: the two grids in this workgroup's slice of HBM scratch (kept zero between uses:
  // every cell a log adds to is cleared when it is read), the coarse (residual, weight)
  // pairs read from memory; otherwise both in LDS
  const int rs = GMEM ? nb : nb + (USE_O ? MERIT_ROW_PAD : 0), nlds = nb * rs;
  long long *G = GMEM ? A.gscratch + (size_t)blockIdx.x * 2 * nlds : (long long *)(bwc_lds + ncb);
  long long *O = USE_O ? G + nlds : nullptr;
  // GMEM: one bit per level L-1 node (2 x 2 cells) that the current orbit's log (maskO) or
  // the trial's log (maskG) touched; the pyramid and the delta pass read only those nodes' cells
  const int nmw = GMEM ? (ngrid / 4 + 31) / 32 : 0;
  unsigned int *maskO = GMEM ? (unsigned int *)bwc_lds : nullptr;
  unsigned int *maskG = GMEM ? maskO + nmw : nullptr;
  int *tbl = GMEM ? (int *)(maskG + nmw) : (int *)(G + (USE_O ? 2 : 1) * nlds);
  __shared__ double s_dF[64];
  __shared__ int s_best, s_n[65];

  const int bid = A.block_base + blockIdx.x;
  const int ti = A.trial_base + bid;
  const int j = A.slot_pos ? A.slot_pos[bid] : bid;
  const int p = A.tidx[ti];
  const int t = A.P.type[p];
  const double *S = A.stacks + (size_t)t * ST_COUNT * Nstack;
  const double2 *bw = A.merit_bw + (size_t)t * kFields * Nstack;
  const size_t o0 = (size_t)j * (K + 1);
  const int4 bo = A.bbox[o0];
  const double x = A.P.x[p], y = A.P.y[p], z = A.P.z[p], mass = A.P.mass[p];
  const double dold[4] = {A.P.vr2[p], A.P.vt2[p], A.P.vp2[p], A.P.vq2[p]};
  const double MT = A.MType[t];
  // eval_fit_mod's combination (main.c:1215-1216): D/M + (f_rs V_rs + f_ts V_ts + f_qs V_qs + f_ps V_ps)/M/4
  const double coef[kFields] = {1.0 / MT, A.fac[t][0] / MT / 4, A.fac[t][1] / MT / 4, A.fac[t][2] / MT / 4,
                                A.fac[t][3] / MT / 4};

  // block tables of the K+1 logs, clean grids, zero merit sums
  for(int q = tid; q <= K; q += nt) s_n[q] = A.log_n[o0 + q];
  for(int q = tid; q < MERIT_MAX_WAVES * K; q += nt) part[q] = 0;
  __syncthreads();
  for(int q = tid; q < (K + 1) * A.log_M; q += nt) {
    const int o = q / A.log_M, b = q % A.log_M;
    if(b * LOG_B < s_n[o]) tbl[q] = A.log_blocks[(o0 + o) * A.log_M + b];
  }
  if(!GMEM)
    for(int q = tid; q < (USE_O ? 2 : 1) * nlds; q += nt) G[q] = 0;
  for(int q = tid; q < 2 * nmw; q += nt) maskO[q] = 0;
  __syncthreads();

  LogRegs R;
  for(int f = 0; f < kFields; f++) {
    const double2 *bwc = GMEM ? bw + (size_t)f * Nstack : bwc_lds;
    if(!GMEM)
      for(int q = tid; q < ncb; q += nt) bwc_lds[q] = bw[(size_t)f * Nstack + q];
    if(USE_O) {
      log_load(A, tbl, s_n[0], f, R);
      log_add(A, tbl, s_n[0], f, R, nb, L, rs, 1.0f, (unsigned long long *)O, maskO);
    }
    // each trial's first records are loaded before the previous trial's pyramid,
    // so their latency overlaps it
    if(K > 0) log_load(A, tbl + A.log_M, s_n[1], f, R);
    for(int k = 0; k < K; k++) {
      const int4 bn = A.bbox[o0 + 1 + k];
      const int4 box = make_int4(min(bo.x, bn.x), max(bo.y, bn.y), min(bo.z, bn.z), max(bo.w, bn.w));
      log_add(A, tbl + (1 + k) * A.log_M, s_n[1 + k], f, R, nb, L, rs, 1.0f, (unsigned long long *)G, maskG);
      if(!USE_O) {
        log_load(A, tbl, s_n[0], f, R);
        log_add(A, tbl, s_n[0], f, R, nb, L, rs, -1.0f, (unsigned long long *)G);
      }
      if(k + 1 < K) log_load(A, tbl + (2 + k) * A.log_M, s_n[2 + k], f, R);
      __syncthreads();
      double acc = merit_block(bw + (size_t)f * Nstack, bwc, L, rs, A.log_unit[f], G, O, box, topb, maskO, maskG);
      __syncthreads();
      for(int q = tid; q < nmw; q += nt) maskG[q] = 0;  // the next trial's marks come after merit_top_levels' barriers
      if(GMEM)
        acc += merit_top_levels(bwc, lblk, A.log_unit[f], top);
      else
        acc += merit_top(bwc, L, A.log_unit[f], topb);
      acc = wave_sum(coef[f] * acc);
      if(lane == 0) part[k * MERIT_MAX_WAVES + wave] += acc;
    }
    if(GMEM)
      clear_marked(O, nb, maskO, nmw);
    else if(USE_O)
      clear_box(O, rs, bo);
    __syncthreads();
    for(int q = tid; q < nmw; q += nt) maskO[q] = 0;
    __syncthreads();
  }
  const int best = choose_trial(M, A, ti, p, t, o0, part, s_dF, &s_best);
  if(best >= 0 && !A.no_deltas) {
    // the orbit-averaged deltas of the taken trial, exact in fixed point
    long long *lad = A.locA_dq + (size_t)t * kFields * ngrid;
    const int4 bn = A.bbox[o0 + 1 + best];
    const int r0 = min(bo.x, bn.x), r1 = max(bo.y, bn.y), z0 = min(bo.z, bn.z), z1 = max(bo.w, bn.w);
    const int nr = (r1 < r0) ? 0 : r1 - r0 + 1, nzz = (z1 < z0) ? 0 : z1 - z0 + 1;
    for(int f = 0; f < kFields; f++) {
      log_load(A, tbl + (1 + best) * A.log_M, s_n[1 + best], f, R);
      log_add(A, tbl + (1 + best) * A.log_M, s_n[1 + best], f, R, nb, L, rs, 1.0f, (unsigned long long *)G, maskG);
      log_load(A, tbl, s_n[0], f, R);
      log_add(A, tbl, s_n[0], f, R, nb, L, rs, -1.0f, (unsigned long long *)G, maskG);
      __syncthreads();
      if(GMEM) {
        // the four cells of each marked node, a thread per mask word
        const int nbp = nb >> 1;
        for(int q = tid; q < nmw; q += nt) {
          for(unsigned int bits = maskG[q]; bits; bits &= bits - 1) {
            const int nd = q * 32 + __builtin_ctz(bits), iz = 2 * (nd / nbp), iR = 2 * (nd % nbp);
#pragma unroll
            for(int c = 0; c < 4; c++) {
              const int cell = (iz + (c >> 1)) * nb + iR + (c & 1);
              const long long d = G[cell];
              G[cell] = 0;
              if(d) atomic_q(&lad[(size_t)f * ngrid + cell], (double)d * A.log_to_loc[f]);
            }
          }
        }
        __syncthreads();
        for(int q = tid; q < nmw; q += nt) maskG[q] = 0;
      } else {
        for(int q = tid; q < nr * nzz; q += nt) {
          const int cell = (z0 + q / nr) * nb + (r0 + q % nr), lc = (z0 + q / nr) * rs + (r0 + q % nr);
          const long long d = G[lc];
          G[lc] = 0;
          if(d) atomic_q(&lad[(size_t)f * ngrid + cell], (double)d * A.log_to_loc[f]);
        }
      }
      __syncthreads();
    }
  }
}

template __global__ void k_eval_log32<true, false, 0>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_log32<true, false, 6>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_log32<false, false, 0>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_log32<false, false, 7>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_log32<true, true, 0>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_log32<true, true, 8>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_log32<true, true, 9>(DevModel M, EvalArgs32 A);

// ----------------------------------------------------- FP32 merit in tiles --
//
// The same merit change as k_eval_log32, for grids whose two int64 copies do not fit in
// LDS (256^2 and 512^2 cells, or 128^2 with long block tables).  Per particle:
//   1. the records of its K+1 logs are listed by tile (a record whose 2x2 cells straddle a
//      tile edge is listed in each tile it touches): a count per tile, offsets, then the
//      entries (log << 24 | record) in this workgroup's slice of HBM;
//   2. field by field, the tiles the logs touch are taken in turn: every listed record's
//      cells inside the tile are added to its log's tile grid with LDS integer atomics (so
//      D = G - O is exact), then each (16 x 16 block, trial) pair is one wave's pyramid from
//      the cells up to the block node, with the terms of every node the difference reaches
//      (cells, 2x2 nodes, 4x4, 8x8, the block); the block sums make the tile's node;
//   3. the levels above the tiles, one wave per trial;
//   4. the choice of the best trial and, if one is taken, its deltas tile by tile.
// The orbit's own log O is summed into one of two grids that alternate between tiles, so the
// grid of the previous tile is cleared while the current one is evaluated: two barriers per
// tile.  Trial grids are cleared as they are read.

// Tiles touched by the 2x2 cells of quad (cells (iz..iz+1, iR..iR+1)); ts tiles per side.
template <typename Visit>
__device__ __forceinline__ void tile_visit(int quad, int L, int ts, Visit visit) {
  const int iz = quad >> L, iR = quad & ((1 << L) - 1);
  const int z0 = iz / TILE, z1 = (iz + 1) / TILE, r0 = iR / TILE, r1 = (iR + 1) / TILE;
  visit(z0 * ts + r0);
  if(r1 != r0) visit(z0 * ts + r1);
  if(z1 != z0) {
    visit(z1 * ts + r0);
    if(r1 != r0) visit(z1 * ts + r1);
  }
}

// Offsets of the tiles' lists (exclusive scan of the counts, off[ntile] = total) and the
// tiles with entries in increasing order; one wave.
__device__ __forceinline__ void tile_scan(const int *cnt, int *off, int *act, int *nact, int ntile, int lane) {
  const int per = (ntile + 63) / 64, a = min(lane * per, ntile), e = min(a + per, ntile);
  int s = 0, na = 0;
  for(int i = a; i < e; i++) {
    s += cnt[i];
    na += cnt[i] > 0;
  }
  int si = s, ni = na;
#pragma unroll
  for(int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(si, d), z = __shfl_up(ni, d);
    if(lane >= d) {
      si += y;
      ni += z;
    }
  }
  int base = si - s, abase = ni - na;
  for(int i = a; i < e; i++) {
    off[i] = base;
    base += cnt[i];
    if(cnt[i] > 0) act[abase++] = i;
  }
  if(lane == 63) {
    off[ntile] = si;
    *nact = ni;
  }
}

// Adds the cells of a record that fall inside the tile (tile-local first cell (lz, lr)).
__device__ __forceinline__ void tile_add(unsigned long long *g, int lz, int lr, float4 x, float sc) {
  const float v[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
  for(int c = 0; c < 4; c++) {
    const int cz = lz + (c >> 1), cr = lr + (c & 1);
    if((unsigned)cz < (unsigned)TILE && (unsigned)cr < (unsigned)TILE && v[c] != 0.f)
      atomicAdd(&g[cz * TILE_RS + cr], (unsigned long long)__float2ll_rn(v[c] * sc));
  }
}

// Index of a node of a tile's pyramid in the tile-local array of (residual, weight) pairs:
// level L - u (u = 0 .. 4), tile-local row i and column j of that level.
__device__ __forceinline__ int tile_node(int u, int i, int j) {
  const int base[5] = {0, 1024, 1280, 1344, 1360};
  return base[u] + i * (TILE >> u) + j;
}

// The tile's (residual, weight) pairs of levels L .. L-4 into LDS (all threads).
__device__ __forceinline__ void tile_load_bw(const double2 *bwf, int L, int tz, int tr, double2 *tbw) {
  for(int q = threadIdx.x; q < TILE_NODES; q += blockDim.x) {
    int u = 0, r = q;
    while(u < 4 && r >= (TILE >> u) * (TILE >> u)) {
      r -= (TILE >> u) * (TILE >> u);
      u++;
    }
    const int n = TILE >> u, i = r / n, j = r % n;
    tbw[q] = bwf[GALIC_STACKOFFSET(L - u, tz * n + i, tr * n + j)];
  }
}

// One wave's pyramid of one 16 x 16 block b (0..3, row-major) of tile (tz, tr) for one trial:
// D = G - O on the block's cells (G cleared as read), the terms of the cells, the 2x2 nodes
// (one per lane, lanes in Morton order) and of the 4x4, 8x8 and 16x16 nodes formed with
// group_sum_i64.  The block node's sum goes to *bs (lane 0).  Returns the lane's terms.
// The (residual, weight) pairs come from the tile's LDS copy (tbw) or from memory (bwf).
template <bool BWL>
__device__ __forceinline__ double tile_block(const double2 *bwf, const double2 *tbw, int L, double unit, long long *G,
                                             const long long *O, int tz, int tr, int b, long long *bs) {
  const int nb = 1 << L, lane = threadIdx.x & 63;
  const int jx = morton_x(lane), jy = morton_y(lane);
  const int lz = (b >> 1) * 16 + 2 * jy, lr = (b & 1) * 16 + 2 * jx;
  const int gz = tz * TILE + lz, gr = tr * TILE + lr;
  long long d[4], s = 0;
  bool nz = false;
#pragma unroll
  for(int c = 0; c < 4; c++) {
    const int q = (lz + (c >> 1)) * TILE_RS + lr + (c & 1);
    d[c] = G[q] - O[q];
    G[q] = 0;
    s += d[c];
    nz = nz || d[c] != 0;
  }
  double acc = 0;
  if(__ballot(nz) == 0) {
    if(lane == 0) *bs = 0;
    return 0.0;
  }
  if(nz) {
    const int offL = GALIC_STACKOFFSET(L, 0, 0);
#pragma unroll
    for(int c = 0; c < 4; c++)
      if(d[c] != 0)
        acc += term_at(BWL ? tbw[tile_node(0, lz + (c >> 1), lr + (c & 1))]
                           : bwf[offL + (gz + (c >> 1)) * nb + gr + (c & 1)],
                       (double)d[c] * unit);
    if(s != 0)
      acc += term_at(BWL ? tbw[tile_node(1, lz >> 1, lr >> 1)] : bwf[GALIC_STACKOFFSET(L - 1, gz >> 1, gr >> 1)],
                     (double)s * unit);
  }
#pragma unroll
  for(int u = 0; u < 3; u++) {
    s = group_sum_i64(s, u);
    const int l = L - 2 - u, mask = (1 << (2 * u + 2)) - 1, sh = L - l;
    if((lane & mask) == 0 && s != 0)
      acc += term_at(BWL ? tbw[tile_node(u + 2, lz >> sh, lr >> sh)] : bwf[GALIC_STACKOFFSET(l, gz >> sh, gr >> sh)],
                     (double)s * unit);
  }
  if(lane == 0) *bs = s;
  return acc;
}

// Largest total record count of one particle's K+1 logs in a launch of n particles.
__global__ void k_log_totals(int n, int K, const int *slot_pos, const int *log_n, unsigned int *maxsum) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if(b >= n) return;
  const size_t o0 = (size_t)(slot_pos ? slot_pos[b] : b) * (K + 1);
  unsigned int s = 0;
  for(int o = 0; o <= K; o++) s += (unsigned int)max(log_n[o0 + o], 0);
  atomicMax(maxsum, s);
}

template <int LF, bool BWL>
__global__ void __launch_bounds__(1024) k_eval_tile32(DevModel M, EvalArgs32 A) {
  extern __shared__ double lds_d[];
  const int tid = threadIdx.x, nt = blockDim.x, K = A.K, wave = tid >> 6, lane = tid & 63, nw = nt >> 6;
  const int L = LF ? LF : M.DG_L, nb = 1 << L, ngrid = nb * nb, Nstack = GALIC_NSTACK(L);
  const int lt = L - 5, ts = 1 << lt, ntile = ts * ts, nst = GALIC_NSTACK(lt);
  double *part = lds_d;                                          // [K][MERIT_MAX_WAVES]
  long long *topst = (long long *)(part + MERIT_MAX_WAVES * K);   // [K][nst]: levels 0..lt per trial
  long long *bsum = topst + (size_t)K * nst;                      // [2][K][4] block sums, alternating
  long long *grids = bsum + 8 * K;                                // O (two, alternating), then G_1..G_K
  unsigned long long *tmask = (unsigned long long *)(grids + (size_t)(K + 2) * TILE_CELLS);  // [ntile] logs touching
  int *tcount = (int *)(tmask + ntile);                           // [ntile] counts, then cursors
  int *toff = tcount + ntile;                                     // [ntile + 1]
  int *tact = toff + ntile + 1;                                   // [ntile] tiles with entries
  int *tbl = tact + ntile;                                        // [K+1][log_M]
  // (offsets from the LDS base keep the pointers in the LDS address space: ds_ instructions)
  const size_t tbw_off = (((size_t)((char *)(tbl + (K + 1) * A.log_M) - (char *)lds_d)) + 15) & ~(size_t)15;
  double2 *tbw = (double2 *)((char *)lds_d + tbw_off);  // BWL: [TILE_NODES]
  __shared__ double s_dF[64];
  __shared__ int s_n[TILE_MAX_K + 1], s_nact, s_best, s_over;
  unsigned int *list = A.tlist + (size_t)blockIdx.x * A.tlist_cap;

  for(int q = tid; q < (K + 2) * TILE_CELLS; q += nt) grids[q] = 0;
  for(int bid = blockIdx.x; bid < A.nblk; bid += gridDim.x) {
    const int ti = A.trial_base + bid;
    const int j = A.slot_pos ? A.slot_pos[bid] : bid;
    const int p = A.tidx[ti];
    const int t = A.P.type[p];
    const double2 *bw = A.merit_bw + (size_t)t * kFields * Nstack;
    const size_t o0 = (size_t)j * (K + 1);
    const double MT = A.MType[t];
    // eval_fit_mod's combination (main.c:1215-1216): D/M + (f_rs V_rs + f_ts V_ts + f_qs V_qs + f_ps V_ps)/M/4
    // (coef_of(f): the weight of field f's terms; a function, not an array indexed at run time)
    auto coef_of = [&](int f) { return f == 0 ? 1.0 / MT : A.fac[t][f - 1] / MT / 4; };
    __syncthreads();  // the previous particle's use of the shared arrays is over
    for(int q = tid; q <= K; q += nt) s_n[q] = max(A.log_n[o0 + q], 0);
    for(int q = tid; q < MERIT_MAX_WAVES * K; q += nt) part[q] = 0;
    for(int q = tid; q < K * nst; q += nt) topst[q] = 0;
    for(int q = tid; q < ntile; q += nt) {
      tcount[q] = 0;
      tmask[q] = 0;
    }
    __syncthreads();
    for(int q = tid; q < (K + 1) * A.log_M; q += nt) {
      const int o = q / A.log_M, b = q % A.log_M;
      if(b * LOG_B < s_n[o]) tbl[q] = A.log_blocks[(o0 + o) * A.log_M + b];
    }
    __syncthreads();
    // 1. the records' tile lists
    for(int o = 0; o <= K; o++) {
      const int n = s_n[o], *tb = tbl + o * A.log_M;
      for(int r = tid; r < n; r += nt) {
        const int quad = A.log_q[(size_t)tb[r / LOG_B] * LOG_B + (r & (LOG_B - 1))];
        tile_visit(quad, L, ts, [&](int tile) {
          atomicAdd(&tcount[tile], 1);
          if(!(tmask[tile] & (1ull << o))) atomicOr(&tmask[tile], 1ull << o);
        });
      }
    }
    __syncthreads();
    if(wave == 0) tile_scan(tcount, toff, tact, &s_nact, ntile, lane);
    __syncthreads();
    if(toff[ntile] > A.tlist_cap) {
      // cannot happen when the caller sizes tlist_cap from k_log_totals; reported, not evaluated
      if(tid == 0) {
        atomicOr(A.tile_err, 1u);
        A.accept[ti] = 0;
        A.dF[ti] = 0;
      }
      continue;
    }
    for(int q = tid; q < ntile; q += nt) tcount[q] = 0;
    __syncthreads();
    for(int o = 0; o <= K; o++) {
      const int n = s_n[o], *tb = tbl + o * A.log_M;
      for(int r = tid; r < n; r += nt) {
        const int quad = A.log_q[(size_t)tb[r / LOG_B] * LOG_B + (r & (LOG_B - 1))];
        tile_visit(quad, L, ts, [&](int tile) {
          list[toff[tile] + atomicAdd(&tcount[tile], 1)] = ((unsigned int)o << 24) | (unsigned int)r;
        });
      }
    }
    __syncthreads();
    const int nact = s_nact;
    // 2. field by field, tile by tile.  The entries (and the tile's pairs) of the next
    // (field, tile) step are loaded into registers while the current tile's pyramid runs, so
    // their latency is hidden; the step after the barrier adds them from registers.
    constexpr int PF = 2;
    int pf_o[PF], pf_quad[PF];
    float4 pf_x[PF];
    double2 pf_bw[2];
    auto prefetch = [&](int step) {
#pragma unroll
      for(int u = 0; u < PF; u++) pf_o[u] = -1;
      if(step >= kFields * nact) return;
      const int f2 = step / nact, tile = tact[step % nact];
#pragma unroll
      for(int u = 0; u < PF; u++) {
        const int e = toff[tile] + tid + u * nt;
        if(e < toff[tile + 1]) {
          const unsigned int en = list[e];
          const int o = (int)(en >> 24), r = (int)(en & 0xffffffu);
          const int blk = tbl[o * A.log_M + r / LOG_B], jj = r & (LOG_B - 1);
          pf_o[u] = o;
          pf_quad[u] = A.log_q[(size_t)blk * LOG_B + jj];
          pf_x[u] = A.log_v[((size_t)blk * LOG_B + jj) * kFields + f2];
        }
      }
      if(BWL) {
        const double2 *bwn = bw + (size_t)f2 * Nstack;
        const int tz = tile / ts, tr = tile % ts;
#pragma unroll
        for(int u = 0; u < 2; u++) {
          const int q = tid + u * nt;
          if(q < TILE_NODES) {
            int lv = 0, r = q;
            while(lv < 4 && r >= (TILE >> lv) * (TILE >> lv)) {
              r -= (TILE >> lv) * (TILE >> lv);
              lv++;
            }
            const int n = TILE >> lv;
            pf_bw[u] = bwn[GALIC_STACKOFFSET(L - lv, tz * n + r / n, tr * n + r % n)];
          }
        }
      }
    };
    prefetch(0);
    for(int f = 0; f < kFields; f++) {
      const float sc = A.log_scale[f];
      const double unit = A.log_unit[f];
      const double2 *bwf = bw + (size_t)f * Nstack;
      for(int ai = 0; ai <= nact; ai++) {
        const int cur = ai & 1;
        if(ai < nact) {
          const int tile = tact[ai], tz = tile / ts, tr = tile % ts;
#pragma unroll
          for(int u = 0; u < PF; u++)
            if(pf_o[u] >= 0) {
              long long *g = grids + (size_t)(pf_o[u] == 0 ? cur : pf_o[u] + 1) * TILE_CELLS;
              tile_add((unsigned long long *)g, (pf_quad[u] >> L) - tz * TILE, (pf_quad[u] & (nb - 1)) - tr * TILE,
                       pf_x[u], sc);
            }
          for(int e = toff[tile] + tid + PF * nt; e < toff[tile + 1]; e += nt) {
            const unsigned int en = list[e];
            const int o = (int)(en >> 24), r = (int)(en & 0xffffffu);
            const int blk = tbl[o * A.log_M + r / LOG_B], jj = r & (LOG_B - 1);
            const int quad = A.log_q[(size_t)blk * LOG_B + jj];
            const float4 x = A.log_v[((size_t)blk * LOG_B + jj) * kFields + f];
            long long *g = grids + (size_t)(o == 0 ? cur : o + 1) * TILE_CELLS;
            tile_add((unsigned long long *)g, (quad >> L) - tz * TILE, (quad & (nb - 1)) - tr * TILE, x, sc);
          }
          if(BWL)
#pragma unroll
            for(int u = 0; u < 2; u++)
              if(tid + u * nt < TILE_NODES) tbw[tid + u * nt] = pf_bw[u];
        }
        __syncthreads();
        if(ai > 0) {
          // the previous tile: its node's sums from the block sums, its O grid cleared
          const int ptile = tact[ai - 1], pz = ptile / ts, pr = ptile % ts;
          const long long *bs = bsum + (size_t)(1 - cur) * 4 * K;
          for(int k = tid; k < K; k += nt)
            topst[(size_t)k * nst + GALIC_STACKOFFSET(lt, pz, pr)] = bs[4 * k] + bs[4 * k + 1] + bs[4 * k + 2] + bs[4 * k + 3];
          long long *Op = grids + (size_t)(1 - cur) * TILE_CELLS;
          for(int q = tid; q < TILE * TILE; q += nt) Op[(q / TILE) * TILE_RS + q % TILE] = 0;
        }
        if(ai < nact) {
          const int tile = tact[ai], tz = tile / ts, tr = tile % ts;
          prefetch(f * nact + ai + 1);
          const long long *O = grids + (size_t)cur * TILE_CELLS;
          const unsigned long long tm = tmask[tile];
          for(int it = wave; it < 4 * K; it += nw) {
            const int k = it >> 2, b = it & 3;
            long long *bs = bsum + (size_t)cur * 4 * K + 4 * k + b;
            if(!(tm & (1ull | (2ull << k)))) {
              // neither the orbit nor the trial reaches this tile: D = 0 (the trial's grid is clean)
              if(lane == 0) *bs = 0;
              continue;
            }
            double acc = tile_block<BWL>(bwf, tbw, L, unit, grids + (size_t)(k + 2) * TILE_CELLS, O, tz, tr, b, bs);
            acc = wave_sum(coef_of(f) * acc);
            if(lane == 0) part[k * MERIT_MAX_WAVES + wave] += acc;
          }
        }
        __syncthreads();
      }
      // 3. the levels above the tiles, one wave per trial
      for(int k = wave; k < K; k += nw) {
        long long *tk = topst + (size_t)k * nst;
        double acc = 0;
        for(int l = lt; l >= 0; l--) {
          const int n = 1 << l;
          for(int q = lane; q < n * n; q += 64) {
            const int i = q >> l, jn = q & (n - 1);
            long long s;
            if(l == lt) {
              s = tk[GALIC_STACKOFFSET(l, i, jn)];
            } else {
              s = tk[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * jn)] + tk[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * jn + 1)] +
                  tk[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * jn)] + tk[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * jn + 1)];
              tk[GALIC_STACKOFFSET(l, i, jn)] = s;
            }
            if(s != 0) acc += term_at(bwf[GALIC_STACKOFFSET(l, i, jn)], (double)s * unit);
          }
          // the next level reads what other lanes of this wave wrote
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        acc = wave_sum(coef_of(f) * acc);
        if(lane == 0) part[k * MERIT_MAX_WAVES + wave] += acc;
      }
      __syncthreads();
    }
    // 4. the choice and the deltas of the taken trial, exact in fixed point
    const int best = choose_trial(M, A, ti, p, t, o0, part, s_dF, &s_best);
    if(best >= 0 && !A.no_deltas) {
      long long *lad = A.locA_dq + (size_t)t * kFields * ngrid;
      const unsigned int ob = (unsigned int)best + 1;
      long long *G = grids;  // O's first grid, zero here
      for(int f = 0; f < kFields; f++) {
        const float sc = A.log_scale[f];
        for(int ai = 0; ai < nact; ai++) {
          const int tile = tact[ai], tz = tile / ts, tr = tile % ts;
          if(!(tmask[tile] & (1ull | (1ull << ob)))) continue;
          for(int e = toff[tile] + tid; e < toff[tile + 1]; e += nt) {
            const unsigned int en = list[e], o = en >> 24;
            if(o != 0 && o != ob) continue;
            const int r = (int)(en & 0xffffffu);
            const int blk = tbl[o * A.log_M + r / LOG_B], jj = r & (LOG_B - 1);
            const int quad = A.log_q[(size_t)blk * LOG_B + jj];
            const float4 x = A.log_v[((size_t)blk * LOG_B + jj) * kFields + f];
            tile_add((unsigned long long *)G, (quad >> L) - tz * TILE, (quad & (nb - 1)) - tr * TILE, x,
                     o == 0 ? -sc : sc);
          }
          __syncthreads();
          for(int q = tid; q < TILE * TILE; q += nt) {
            const int lc = (q / TILE) * TILE_RS + q % TILE;
            const long long d = G[lc];
            if(d) {
              G[lc] = 0;
              atomic_q(&lad[(size_t)f * ngrid + (size_t)(tz * TILE + q / TILE) * nb + tr * TILE + q % TILE],
                       (double)d * A.log_to_loc[f]);
            }
          }
          __syncthreads();
        }
      }
    }
  }
}

template __global__ void k_eval_tile32<0, false>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_tile32<0, true>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_tile32<9, false>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_tile32<9, true>(DevModel M, EvalArgs32 A);

// ------------------------------------- trial choice around the tile merit --
//
// The tile merit kernels with the instantaneous terms and the write-back of choose_trial
// split off into two small kernels, which keeps the FP64 moment code out of the merit
// kernel's registers.

// The instantaneous merit change of every trial of the launch (inst_delta_q), one thread
// per trial: vinst[ti * K + k].
__global__ void k_trial_inst(DevModel M, EvalArgs32 A, double *vinst) {
  const int K = A.K;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if(q >= A.nblk * K) return;
  const int ti = A.trial_base + q / K, k = q % K;
  const int p = A.tidx[ti], t = A.P.type[p];
  const int nb = M.EG_Nbin, ngrid = nb * nb, Nstack = GALIC_NSTACK(M.DG_L);
  const double *S = A.stacks + (size_t)t * ST_COUNT * Nstack;
  const double x = A.P.x[p], y = A.P.y[p], z = A.P.z[p], mass = A.P.mass[p];
  const double dold[4] = {A.P.vr2[p], A.P.vt2[p], A.P.vp2[p], A.P.vq2[p]};
  const size_t qq = (size_t)ti * K + k;
  double dnew[4];
  disp_components64(M, t, x, y, z, A.tvx[qq], A.tvy[qq], A.tvz[qq], dnew[0], dnew[1], dnew[2], dnew[3]);
  double cw[4];
  int cc[4];
  vinst[qq] = inst_delta_q(M, S, A.cutB_of + (size_t)t * ngrid, t, x, y, z, mass, dnew, dold, A.lowdisp, cw, cc);
}

// The taken trial of each particle of the launch (A.accept = trial + 1, or 0), one thread per
// particle: orbit count and step count of the kept orbit; for a taken trial its
// instantaneous deltas and the particle's new velocity and moments (main.c:480-527).
__global__ void k_apply_best(DevModel M, EvalArgs32 A) {
  const int bid = blockIdx.x * blockDim.x + threadIdx.x;
  if(bid >= A.nblk) return;
  const int K = A.K, ti = A.trial_base + bid;
  const int j = A.slot_pos ? A.slot_pos[bid] : bid;
  const int p = A.tidx[ti], t = A.P.type[p];
  const int best = A.accept[ti] - 1;
  const size_t o0 = (size_t)j * (K + 1), oc = (best >= 0) ? o0 + 1 + best : o0;
  A.P.orbits[p] = A.norbit[oc];
  A.psteps[p] = (int)min(A.osteps[oc], (long long)0x7fffffff);
  if(best < 0) return;
  const int nb = M.EG_Nbin, ngrid = nb * nb;
  const double x = A.P.x[p], y = A.P.y[p], z = A.P.z[p], mass = A.P.mass[p];
  const double dold[4] = {A.P.vr2[p], A.P.vt2[p], A.P.vp2[p], A.P.vq2[p]};
  const size_t q = (size_t)ti * K + best;
  const double vt[3] = {A.tvx[q], A.tvy[q], A.tvz[q]};
  double dnew[4];
  disp_components64(M, t, x, y, z, vt[0], vt[1], vt[2], dnew[0], dnew[1], dnew[2], dnew[3]);
  int iR, iz;
  double fR, fz;
  cell64(x, y, z, M.EG_Rmin, M.EG_logFac, M.EG_Nbin, iR, iz, fR, fz);
  const double cw[4] = {(1 - fR) * (1 - fz), fR * (1 - fz), (1 - fR) * fz, fR * fz};
  const int cc[4] = {iz * nb + iR, iz * nb + iR + 1, (iz + 1) * nb + iR, (iz + 1) * nb + iR + 1};
  long long *lid = A.locI_dq + (size_t)t * 4 * ngrid;
  for(int f = 0; f < 4; f++) {
    const double v = (dnew[f] - dold[f]) * mass;
    for(int c = 0; c < 4; c++) atomic_q(&lid[(size_t)f * ngrid + cc[c]], A.iscale * cw[c] * v);
  }
  A.P.vx[p] = vt[0];
  A.P.vy[p] = vt[1];
  A.P.vz[p] = vt[2];
  A.P.vr2[p] = dnew[0];
  A.P.vt2[p] = dnew[1];
  A.P.vp2[p] = dnew[2];
  A.P.vq2[p] = dnew[3];
}

// ------------------------------------------- FP32 merit in sparse tiles --
//
// k_eval_tile32 with the pyramid of each tile restricted to the 2x2 nodes the logs touch.
// The adds of a tile mark the 2x2 nodes of the cells they reach (a bitmap of the tile's
// 256 nodes); the pyramid then takes (node, trial) pairs of marked nodes only: a lane forms
// D = G - O on the node's four cells, the terms of the cells and of the node, and adds the
// node's sum into its 4 x 4-cell parent's sum (LDS integer atomics, exact); one wave per trial
// then forms the tile's three coarser levels from the 8 x 8 parent sums with group_sum_i64.
// That last step of a tile runs while the next tile's records are added (the orbit's grid,
// the bitmap, the parent sums and the coarse pairs alternate between two copies), so a tile
// costs two barriers.  Same sums as k_eval_tile32: the same decisions and deltas.

// A workgroup barrier that waits for the wave's LDS operations only: loads from memory
// issued before it stay in flight across it (their registers are waited on where used).
// The phases it separates exchange data through LDS alone.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// The j-th set bit (0-based) of a 256-bit map held as eight words; -1 if there is none.
__device__ __forceinline__ int select_bit(const unsigned int *w, int j) {
#pragma unroll
  for(int q = 0; q < 8; q++) {
    const int c = __popc(w[q]);
    if(j < c) {
      unsigned int x = w[q];
      for(int u = 0; u < j; u++) x &= x - 1;
      return q * 32 + __ffs(x) - 1;
    }
    j -= c;
  }
  return -1;
}

// Adds the cells of a record that fall inside the tile and marks their 2x2 nodes.
__device__ __forceinline__ void tile_add_mark(unsigned long long *g, unsigned int *bits, int lz, int lr, float4 x,
                                              float sc) {
  const float v[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
  for(int c = 0; c < 4; c++) {
    const int cz = lz + (c >> 1), cr = lr + (c & 1);
    if((unsigned)cz < (unsigned)TILE && (unsigned)cr < (unsigned)TILE && v[c] != 0.f) {
      atomicAdd(&g[cz * TILE_RS + cr], (unsigned long long)__float2ll_rn(v[c] * sc));
      const int nd = (cz >> 1) * (TILE / 2) + (cr >> 1);
      const unsigned int bit = 1u << (nd & 31);
      if(!(bits[nd >> 5] & bit)) atomicOr(&bits[nd >> 5], bit);
    }
  }
}

template <int LF>
__global__ void __launch_bounds__(1024) k_eval_sparse32(DevModel M, EvalArgs32 A) {
  extern __shared__ double lds_d[];
  const int tid = threadIdx.x, nt = blockDim.x, K = A.K, wave = tid >> 6, lane = tid & 63, nw = nt >> 6;
  const int L = LF ? LF : M.DG_L, nb = 1 << L, ngrid = nb * nb, Nstack = GALIC_NSTACK(L);
  const int lt = L - 5, ts = 1 << lt, ntile = ts * ts, nst = GALIC_NSTACK(lt);
  double *part = lds_d;                                           // [K][MERIT_MAX_WAVES]
  long long *topst = (long long *)(part + MERIT_MAX_WAVES * K);    // [K][nst]: levels 0..lt per trial
  long long *acc2 = topst + (size_t)K * nst;                       // [2][K][64] sums of the 4x4-cell nodes
  long long *grids = acc2 + 128 * (size_t)K;                       // O (two, alternating), then G_1..G_K
  unsigned long long *tmask = (unsigned long long *)(grids + (size_t)(K + 2) * TILE_CELLS);  // [ntile]
  int *tcount = (int *)(tmask + ntile);                            // [ntile] counts, then cursors
  int *toff = tcount + ntile;                                      // [ntile + 1]
  int *tact = toff + ntile + 1;                                    // [ntile] tiles with entries
  int *tbl_s = tact + ntile;                                       // [K+1][log_M] if staged
  unsigned int *tbits = (unsigned int *)(tbl_s + (A.tbl_lds ? (K + 1) * A.log_M : 0));  // [2][8] marked nodes
  // (offsets from the LDS base keep the pointers in the LDS address space: ds_ instructions)
  const size_t tbw_off = (((size_t)((char *)(tbits + 16) - (char *)lds_d)) + 15) & ~(size_t)15;
  double2 *tbw = (double2 *)((char *)lds_d + tbw_off);            // [1280] cells and 2x2 nodes
  double2 *tbwc = tbw + 1280;                                      // [2][84] 4x4, 8x8, 16x16 nodes
  __shared__ double s_dF[64];
  __shared__ int s_n[TILE_MAX_K + 1], s_nact, s_best;
  // this workgroup's copy of the records, sorted by tile: per entry the quad index with the
  // log in bits 24.. (qo), then per field the entry's four values (val[f])
  const size_t cap = (size_t)A.tlist_cap;
  unsigned int *qo = A.tlist + (size_t)blockIdx.x * cap * 21;
  float4 *val = (float4 *)(qo + cap);

  for(int q = tid; q < (K + 2) * TILE_CELLS; q += nt) grids[q] = 0;
  for(int q = tid; q < 256 * K; q += nt) acc2[q] = 0;
  for(int q = tid; q < 16; q += nt) tbits[q] = 0;
  // optional timing (A.prof): wall clock of the phases of each particle, summed by thread 0
  unsigned long long tp0 = 0, tp1 = 0, tp2 = 0;
  for(int bid = blockIdx.x; bid < A.nblk; bid += gridDim.x) {
    if(A.prof && tid == 0) tp0 = wall_clock64();
    const int ti = A.trial_base + bid;
    const int j = A.slot_pos ? A.slot_pos[bid] : bid;
    const int p = A.tidx[ti];
    const int t = A.P.type[p];
    const double2 *bw = A.merit_bw + (size_t)t * kFields * Nstack;
    const size_t o0 = (size_t)j * (K + 1);
    const double MT = A.MType[t];
    // the block tables of the K+1 logs: in LDS or in memory (rows o0 .. o0+K of log_blocks)
    const int *tbl = A.tbl_lds ? (const int *)tbl_s : A.log_blocks + o0 * A.log_M;
    // eval_fit_mod's combination (main.c:1215-1216): D/M + (f_rs V_rs + f_ts V_ts + f_qs V_qs + f_ps V_ps)/M/4
    auto coef_of = [&](int f) { return f == 0 ? 1.0 / MT : A.fac[t][f - 1] / MT / 4; };
    __syncthreads();  // the previous particle's use of the shared arrays is over
    for(int q = tid; q <= K; q += nt) s_n[q] = max(A.log_n[o0 + q], 0);
    for(int q = tid; q < MERIT_MAX_WAVES * K; q += nt) part[q] = 0;
    for(int q = tid; q < K * nst; q += nt) topst[q] = 0;
    for(int q = tid; q < ntile; q += nt) {
      tcount[q] = 0;
      tmask[q] = 0;
    }
    __syncthreads();
    if(A.tbl_lds)
      for(int q = tid; q < (K + 1) * A.log_M; q += nt) {
        const int o = q / A.log_M, b = q % A.log_M;
        if(b * LOG_B < s_n[o]) tbl_s[q] = A.log_blocks[(o0 + o) * A.log_M + b];
      }
    __syncthreads();
    // 1. the records' tile lists (as k_eval_tile32)
    for(int o = 0; o <= K; o++) {
      const int n = s_n[o], *tb = tbl + o * A.log_M;
      for(int r = tid; r < n; r += nt) {
        const int quad = A.log_q[(size_t)tb[r / LOG_B] * LOG_B + (r & (LOG_B - 1))];
        tile_visit(quad, L, ts, [&](int tile) {
          atomicAdd(&tcount[tile], 1);
          if(!(tmask[tile] & (1ull << o))) atomicOr(&tmask[tile], 1ull << o);
        });
      }
    }
    __syncthreads();
    if(wave == 0) tile_scan(tcount, toff, tact, &s_nact, ntile, lane);
    __syncthreads();
    if(toff[ntile] > A.tlist_cap) {
      // cannot happen when the caller sizes tlist_cap from k_log_totals; reported, not evaluated
      if(tid == 0) {
        atomicOr(A.tile_err, 1u);
        A.accept[ti] = 0;
        A.dF[ti] = 0;
      }
      continue;
    }
    for(int q = tid; q < ntile; q += nt) tcount[q] = 0;
    __syncthreads();
    // the records copied into the tile order, all five fields at once: every later read of a
    // tile's records is one contiguous load per field
    for(int o = 0; o <= K; o++) {
      const int n = s_n[o], *tb = tbl + o * A.log_M;
      for(int r = tid; r < n; r += nt) {
        const int blk = tb[r / LOG_B], jj = r & (LOG_B - 1);
        const int quad = A.log_q[(size_t)blk * LOG_B + jj];
        float4 x[kFields];
#pragma unroll
        for(int f = 0; f < kFields; f++) x[f] = A.log_v[((size_t)blk * LOG_B + jj) * kFields + f];
        // the (one to four) tiles of the record's cells, as tile_visit takes them
        const int iz = quad >> L, iR = quad & (nb - 1);
        const int z0 = iz / TILE, z1 = (iz + 1) / TILE, r0 = iR / TILE, r1 = (iR + 1) / TILE;
        const int tl[4] = {z0 * ts + r0, z0 * ts + r1, z1 * ts + r0, z1 * ts + r1};
        const bool use[4] = {true, r1 != r0, z1 != z0, z1 != z0 && r1 != r0};
#pragma unroll
        for(int u = 0; u < 4; u++)
          if(use[u]) {
            const size_t e = toff[tl[u]] + atomicAdd(&tcount[tl[u]], 1);
            qo[e] = ((unsigned int)o << 24) | (unsigned int)quad;
#pragma unroll
            for(int f = 0; f < kFields; f++) val[f * cap + e] = x[f];
          }
      }
    }
    __syncthreads();
    const int nact = s_nact;
    if(A.prof && tid == 0) {
      tp1 = wall_clock64();
      atomicAdd(&A.prof[0], tp1 - tp0);
      atomicAdd(&A.prof[5], (unsigned long long)nact);
      atomicAdd(&A.prof[6], (unsigned long long)toff[ntile]);
      atomicAdd(&A.prof[7], 1ull);
    }
    // 2. field by field, tile by tile; the next step's entries and pairs are loaded into
    // registers during the current tile's pyramid
    constexpr int PF = 1;
    struct Prefetch {
      int o[PF], quad[PF];
      float4 x[PF];
      double2 bw[2];
    };
    auto prefetch = [&](int step) {
      Prefetch r;
#pragma unroll
      for(int u = 0; u < PF; u++) r.o[u] = -1;
      if(step >= kFields * nact) return r;
      const int f2 = step / nact, tile = tact[step % nact];
#pragma unroll
      for(int u = 0; u < PF; u++) {
        const int e = toff[tile] + tid + u * nt;
        if(e < toff[tile + 1]) {
          const unsigned int en = qo[e];
          r.o[u] = (int)(en >> 24);
          r.quad[u] = (int)(en & 0xffffffu);
          r.x[u] = val[f2 * cap + e];
        }
      }
      const double2 *bwn = bw + (size_t)f2 * Nstack;
      const int tz = tile / ts, tr = tile % ts;
#pragma unroll
      for(int u = 0; u < 2; u++) {
        const int q = tid + u * nt;
        if(q < TILE_NODES) {
          int lv = 0, rr = q;
          while(lv < 4 && rr >= (TILE >> lv) * (TILE >> lv)) {
            rr -= (TILE >> lv) * (TILE >> lv);
            lv++;
          }
          const int n = TILE >> lv;
          r.bw[u] = bwn[GALIC_STACKOFFSET(L - lv, tz * n + rr / n, tr * n + rr % n)];
        }
      }
      return r;
    };
    Prefetch pf = prefetch(0);
    for(int f = 0; f < kFields; f++) {
      const float sc = A.log_scale[f];
      const double unit = A.log_unit[f];
      const double2 *bwf = bw + (size_t)f * Nstack;
      for(int ai = 0; ai <= nact; ai++) {
        const int cur = ai & 1, prv = 1 - cur;
        unsigned long long ta = 0;
        if(A.prof && tid == 0) ta = wall_clock64();
        // phase A: this tile's records added and their nodes marked; the previous tile's
        // coarse levels, its orbit grid cleared
        if(ai < nact) {
          const int tile = tact[ai], tz = tile / ts, tr = tile % ts;
          unsigned long long *O = (unsigned long long *)(grids + (size_t)cur * TILE_CELLS);
          unsigned int *bits = tbits + 8 * cur;
#pragma unroll
          for(int u = 0; u < PF; u++)
            if(pf.o[u] >= 0)
              tile_add_mark(pf.o[u] == 0 ? O : (unsigned long long *)(grids + (size_t)(pf.o[u] + 1) * TILE_CELLS),
                            bits, (pf.quad[u] >> L) - tz * TILE, (pf.quad[u] & (nb - 1)) - tr * TILE, pf.x[u], sc);
          for(int e = toff[tile] + tid + PF * nt; e < toff[tile + 1]; e += nt) {
            const unsigned int en = qo[e];
            const int o = (int)(en >> 24), quad = (int)(en & 0xffffffu);
            const float4 x = val[f * cap + e];
            tile_add_mark(o == 0 ? O : (unsigned long long *)(grids + (size_t)(o + 1) * TILE_CELLS), bits,
                          (quad >> L) - tz * TILE, (quad & (nb - 1)) - tr * TILE, x, sc);
          }
#pragma unroll
          for(int u = 0; u < 2; u++) {
            const int q = tid + u * nt;
            if(q < 1280)
              tbw[q] = pf.bw[u];
            else if(q < TILE_NODES)
              tbwc[84 * cur + q - 1280] = pf.bw[u];
          }
        }
        if(ai > 0) {
          const int ptile = tact[ai - 1], pz = ptile / ts, pr = ptile % ts;
          // coarse levels of the previous tile: wave k takes trial k, lanes the 8x8 nodes of
          // 4x4 cells in Morton order
          for(int k = wave; k < K; k += nw) {
            long long *a2 = acc2 + ((size_t)prv * K + k) * 64;
            const double2 *cw = tbwc + 84 * prv;
            const int mx = morton_x(lane), my = morton_y(lane);
            long long s = a2[my * 8 + mx];
            a2[my * 8 + mx] = 0;
            double acc = 0;
            if(s != 0) acc += term_at(cw[my * 8 + mx], (double)s * unit);
            s = group_sum_i64(s, 0);
            if((lane & 3) == 0 && s != 0) acc += term_at(cw[64 + (my >> 1) * 4 + (mx >> 1)], (double)s * unit);
            s = group_sum_i64(s, 1);
            if((lane & 15) == 0 && s != 0) acc += term_at(cw[80 + (my >> 2) * 2 + (mx >> 2)], (double)s * unit);
            s = group_sum_i64(s, 2);
            if(lane == 0) topst[(size_t)k * nst + GALIC_STACKOFFSET(lt, pz, pr)] = s;
            acc = wave_sum(coef_of(f) * acc);
            if(lane == 0) part[k * MERIT_MAX_WAVES + wave] += acc;
          }
          // the previous tile's orbit grid, at its marked nodes
          const unsigned int *pb = tbits + 8 * prv;
          long long *Op = grids + (size_t)prv * TILE_CELLS;
          for(int q = tid; q < 256; q += nt)
            if((pb[q >> 5] >> (q & 31)) & 1u) {
              const int c0 = (q / 16) * 2 * TILE_RS + (q % 16) * 2;
              Op[c0] = 0;
              Op[c0 + 1] = 0;
              Op[c0 + TILE_RS] = 0;
              Op[c0 + TILE_RS + 1] = 0;
            }
        }
        lds_barrier();
        unsigned long long tb = 0;
        if(A.prof && tid == 0) {
          tb = wall_clock64();
          atomicAdd(&A.prof[9], tb - ta);
        }
        // phase B: the pairs (marked node, trial) of this tile
        if(ai > 0 && tid < 8) tbits[8 * prv + tid] = 0;  // the previous tile's map is read no more
        if(ai < nact) {
          const int tile = tact[ai], tz = tile / ts, tr = tile % ts;
          pf = prefetch(f * nact + ai + 1);
          unsigned long long tpf = 0;
          if(A.prof && tid == 0) {
            tpf = wall_clock64();
            atomicAdd(&A.prof[11], tpf - tb);
          }
          const long long *O = grids + (size_t)cur * TILE_CELLS;
          const unsigned int *bits = tbits + 8 * cur;
          int nmark = 0;
#pragma unroll
          for(int q = 0; q < 8; q++) nmark += __popc(bits[q]);
          const int nch = (nmark + 63) >> 6;
          const unsigned long long tm = tmask[tile];
          for(int u = wave; u < K * nch; u += nw) {
            const int k = u / nch, c = u % nch;
            if(!(tm & (1ull | (2ull << k)))) continue;  // neither the orbit nor the trial reaches the tile
            long long *G = grids + (size_t)(k + 2) * TILE_CELLS;
            double acc = 0;
            const int jn = c * 64 + lane;
            if(jn < nmark) {
              const int nd = select_bit(bits, jn), nz = nd / 16, nr = nd % 16;
              const int c0 = 2 * nz * TILE_RS + 2 * nr;
              const int cq[4] = {c0, c0 + 1, c0 + TILE_RS, c0 + TILE_RS + 1};
              long long s = 0;
#pragma unroll
              for(int cc = 0; cc < 4; cc++) {
                const long long d = G[cq[cc]] - O[cq[cc]];
                G[cq[cc]] = 0;
                s += d;
                if(d != 0) acc += term_at(tbw[tile_node(0, 2 * nz + (cc >> 1), 2 * nr + (cc & 1))], (double)d * unit);
              }
              if(s != 0) {
                acc += term_at(tbw[tile_node(1, nz, nr)], (double)s * unit);
                atomicAdd((unsigned long long *)&acc2[((size_t)cur * K + k) * 64 + (nz >> 1) * 8 + (nr >> 1)],
                          (unsigned long long)s);
              }
            }
            acc = wave_sum(coef_of(f) * acc);
            if(lane == 0) part[k * MERIT_MAX_WAVES + wave] += acc;
          }
          if(A.prof && tid == 0) atomicAdd(&A.prof[12], wall_clock64() - tpf);
        }
        lds_barrier();
        if(A.prof && tid == 0) atomicAdd(&A.prof[10], wall_clock64() - tb);
      }
      if(nact == 0 && tid < 16) tbits[tid] = 0;
      // 3. the levels above the tiles, one wave per trial
      for(int k = wave; k < K; k += nw) {
        long long *tk = topst + (size_t)k * nst;
        double acc = 0;
        for(int l = lt; l >= 0; l--) {
          const int n = 1 << l;
          for(int q = lane; q < n * n; q += 64) {
            const int i = q >> l, jn = q & (n - 1);
            long long s;
            if(l == lt) {
              s = tk[GALIC_STACKOFFSET(l, i, jn)];
            } else {
              s = tk[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * jn)] + tk[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * jn + 1)] +
                  tk[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * jn)] + tk[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * jn + 1)];
              tk[GALIC_STACKOFFSET(l, i, jn)] = s;
            }
            if(s != 0) acc += term_at(bwf[GALIC_STACKOFFSET(l, i, jn)], (double)s * unit);
          }
          // the next level reads what other lanes of this wave wrote
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        acc = wave_sum(coef_of(f) * acc);
        if(lane == 0) part[k * MERIT_MAX_WAVES + wave] += acc;
      }
      __syncthreads();
    }
    // 4. the choice and the deltas of the taken trial, exact in fixed point
    if(A.prof && tid == 0) {
      tp2 = wall_clock64();
      atomicAdd(&A.prof[1], tp2 - tp1);
    }
    // the total merit change of each trial, the best improving one taken (the instantaneous
    // terms from k_trial_inst; the write-back follows in k_apply_best)
    for(int k = tid; k < K; k += nt) {
      const size_t q = (size_t)ti * K + k;
      const double *fit = A.fit + (size_t)t * 16;
      double orbit_terms = 0;
      for(int w = 0; w < MERIT_MAX_WAVES; w++) orbit_terms += part[k * MERIT_MAX_WAVES + w];
      s_dF[k] = orbit_terms + fit[5] / fit[6] * A.vinst[q];
      A.dFk[q] = s_dF[k];
    }
    __syncthreads();
    if(tid == 0) {
      int bk = -1;
      double bv = 0;
      for(int k = 0; k < K; k++)
        if(s_dF[k] < bv) {
          bv = s_dF[k];
          bk = k;
        }
      s_best = bk;
      A.accept[ti] = bk + 1;
      A.dF[ti] = (bk >= 0) ? bv : (K > 0 ? s_dF[0] : 0.0);
    }
    __syncthreads();
    const int best = s_best;
    if(A.prof && tid == 0) {
      tp1 = wall_clock64();
      atomicAdd(&A.prof[2], tp1 - tp2);
      if(best >= 0) atomicAdd(&A.prof[8], 1ull);
    }
    if(best >= 0 && !A.no_deltas) {
      long long *lad = A.locA_dq + (size_t)t * kFields * ngrid;
      const unsigned int ob = (unsigned int)best + 1;
      long long *G = grids;  // O's first grid, zero here
      for(int f = 0; f < kFields; f++) {
        const float sc = A.log_scale[f];
        for(int ai = 0; ai < nact; ai++) {
          const int tile = tact[ai], tz = tile / ts, tr = tile % ts;
          if(!(tmask[tile] & (1ull | (1ull << ob)))) continue;
          for(int e = toff[tile] + tid; e < toff[tile + 1]; e += nt) {
            const unsigned int en = qo[e], o = en >> 24;
            if(o != 0 && o != ob) continue;
            const int quad = (int)(en & 0xffffffu);
            const float4 x = val[f * cap + e];
            tile_add((unsigned long long *)G, (quad >> L) - tz * TILE, (quad & (nb - 1)) - tr * TILE, x,
                     o == 0 ? -sc : sc);
          }
          __syncthreads();
          for(int q = tid; q < TILE * TILE; q += nt) {
            const int lc = (q / TILE) * TILE_RS + q % TILE;
            const long long d = G[lc];
            if(d) {
              G[lc] = 0;
              atomic_q(&lad[(size_t)f * ngrid + (size_t)(tz * TILE + q / TILE) * nb + tr * TILE + q % TILE],
                       (double)d * A.log_to_loc[f]);
            }
          }
          __syncthreads();
        }
      }
    }
    if(A.prof && tid == 0) atomicAdd(&A.prof[3], wall_clock64() - tp1);
  }
}

template __global__ void k_eval_sparse32<0>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_sparse32<9>(DevModel M, EvalArgs32 A);

// ------------------------------------------------ FP32 merit, 16x16 tiles --
//
// The merit change of each trial -- the terms of every node of the pyramid that D = G_k - O
// reaches, as k_eval_log32 -- organised for large grids:
//  * Fixed point per particle: a record value v of field f counts round(v * 2^kp[f]) units,
//    kp[f] = 30 - e with the largest total of the particle's K+1 orbits (k_orbit32_log's
//    log_tot) below 2^e.  Every sum of one orbit's values then stays below 2^30 (+ one unit
//    per record of rounding), so tile grids and node sums are exact int32.
//  * Tiles of 16 x 16 cells, taken in Morton order.  The records of the K+1 logs are listed by
//    tile (arena record index, log index; a record whose 2x2 cells straddle a tile edge is
//    listed in each tile it reaches).  A tile's K+1 grids of all five fields sit in LDS at once,
//    node-major: the four cells of a 2x2 node are one 16-byte word.
//  * Wave f evaluates field f.  Lane j owns the tile's 2x2 node j (lanes in Morton order, so
//    groups of 4 and 16 lanes are the 4x4 and 8x8 nodes and the wave is the tile); its
//    (residual, weight) pairs stay in registers for all trials, and for each trial the lane forms
//    D on its four cells, the node sum, the quad and row sums (DPP) and the tile sum, and adds
//    the terms of every node it leads into a per-trial FP64 accumulator.
//  * Levels above the tiles: a tile's ancestor at level l is complete once the walk reaches a
//    tile with another ancestor at l.  Lane k keeps trial k's partial sums of the open
//    ancestors (a stack of one sum per level) and closes them level by level.
// Then the choice (the instantaneous terms come from k_trial_inst; the write-back follows in
// k_apply_best) and the deltas of the taken trial: exact int64 sums at the log scale on
// blocks of 32 x 32 cells (four consecutive tiles), the same values as k_eval_tile32.

__device__ __forceinline__ int spread6(int x) {
  x &= 63;
  x = (x | (x << 4)) & 0x0F0F;
  x = (x | (x << 2)) & 0x3333;
  x = (x | (x << 1)) & 0x5555;
  return x;
}
__device__ __forceinline__ int compact6(int x) {
  x &= 0x5555;
  x = (x | (x >> 1)) & 0x3333;
  x = (x | (x >> 2)) & 0x0F0F;
  x = (x | (x >> 4)) & 0x00FF;
  return x;
}
// Morton code of tile (tz, tr) (tr in the even bits, as morton_x of a lane) and back.
__device__ __forceinline__ int t16_code(int tz, int tr) { return spread6(tr) | (spread6(tz) << 1); }
__device__ __forceinline__ int t16_z(int m) { return compact6(m >> 1); }
__device__ __forceinline__ int t16_r(int m) { return compact6(m); }
// Word of tile-local cell (cz, cr) in a node-major 16 x 16 grid.
__device__ __forceinline__ int t16_cell(int cz, int cr) {
  return ((spread6(cr >> 1) | (spread6(cz >> 1) << 1)) << 2) | ((cz & 1) << 1) | (cr & 1);
}

// Tiles (Morton codes) reached by the 2x2 cells of quad.
template <typename Visit>
__device__ __forceinline__ void t16_visit(int quad, int L, Visit visit) {
  const int iz = quad >> L, iR = quad & ((1 << L) - 1);
  const int z0 = iz >> 4, z1 = (iz + 1) >> 4, r0 = iR >> 4, r1 = (iR + 1) >> 4;
  visit(t16_code(z0, r0));
  if(r1 != r0) visit(t16_code(z0, r1));
  if(z1 != z0) {
    visit(t16_code(z1, r0));
    if(r1 != r0) visit(t16_code(z1, r1));
  }
}

// acc + w (|b + d| - |b|): a node's term with b and d in the particle's units (d = 0 adds 0).
__device__ __forceinline__ double t16_term(double b, double w, double d, double acc) {
  return fma(w, fabs(b + d) - fabs(b), acc);
}

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

// The logs reaching each tile: one bit per log (bit 0 the current orbit, bit k+1 trial k), 16
// bits per tile for up to 15 trials, 32 above.
template <int KR>
__device__ __forceinline__ unsigned int t16_mask_get(const unsigned int *w, int m) {
  if(KR <= 15) return (w[m >> 1] >> ((m & 1) * 16)) & 0xffffu;
  return w[m];
}
template <int KR>
__device__ __forceinline__ void t16_mask_or(unsigned int *w, int m, unsigned int bits) {
  if(KR <= 15)
    atomicOr(&w[m >> 1], bits << ((m & 1) * 16));
  else
    atomicOr(&w[m], bits);
}

// Offsets of the tiles' lists (exclusive scan of the counts, off[ntile] = total) and the tiles
// with entries in increasing order (16-bit indices); one wave.
__device__ __forceinline__ void t16_scan(const int *cnt, int *off, unsigned short *act, int *nact, int ntile, int lane) {
  const int per = (ntile + 63) / 64, a = min(lane * per, ntile), e = min(a + per, ntile);
  int s = 0, na = 0;
  for(int i = a; i < e; i++) {
    s += cnt[i];
    na += cnt[i] > 0;
  }
  int si = s, ni = na;
#pragma unroll
  for(int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(si, d), z = __shfl_up(ni, d);
    if(lane >= d) {
      si += y;
      ni += z;
    }
  }
  int base = si - s, abase = ni - na;
  for(int i = a; i < e; i++) {
    off[i] = base;
    base += cnt[i];
    if(cnt[i] > 0) act[abase++] = (unsigned short)i;
  }
  if(lane == 63) {
    off[ntile] = si;
    *nact = ni;
  }
}

constexpr int T16_LU = 8;  // records a thread takes at once in the list passes

// Record index of the q-th record of the particle's logs (orbits' records in turn; s_off
// the exclusive prefix of the K+1 counts) and its log.
__device__ __forceinline__ unsigned int t16_record(const EvalArgs32 &A, const int *s_off, int K, size_t o0, int q,
                                                   int &o) {
  o = 0;
  while(o < K && q >= s_off[o + 1]) o++;
  const int r = q - s_off[o];
  const int blk = A.log_blocks[(o0 + o) * A.log_M + r / LOG_B];
  return (unsigned int)blk * LOG_B + (unsigned int)(r & (LOG_B - 1));
}

template <int LF, int KR>
__global__ void __launch_bounds__(64 * kFields, 4) k_eval_t16(DevModel M, EvalArgs32 A) {
  extern __shared__ int lds_i[];
  const int tid = threadIdx.x, nt = blockDim.x, K = A.K, wave = tid >> 6, lane = tid & 63;
  const int L = LF ? LF : M.DG_L, nb = 1 << L, Nstack = GALIC_NSTACK(L);
  const int lt = L - 4, ntile = 1 << (2 * lt);
  const int gwords = t16_grid_words(K, L);
  int *grids = lds_i;                                  // [K+1][5][256] int32: O, then G_k; counts while listing
  int *tcnt = grids;                                   // [ntile] (list building only)
  int *toff = grids + gwords;                          // [ntile + 1]
  unsigned short *tact = (unsigned short *)(toff + ntile + 1);  // [ntile] tiles with entries, Morton order
  unsigned int *tmask = (unsigned int *)(toff + ntile + 1 + (ntile + 1) / 2);  // logs reaching each tile
  const int nmask = t16_mask_words(ntile, KR);
  __shared__ double s_part[kFields][KR];
  __shared__ double s_dF[KR];
  __shared__ int s_off[KR + 2], s_kp[kFields], s_nact, s_best, s_item;
  constexpr int NSTK = (LF ? LF : T16_MAX_LEVEL) - 4;
  // trial k's sums of the open ancestors of field f's walk at level l < lt: s_stk[f][l][k]
  __shared__ int s_stk[kFields][NSTK > 0 ? NSTK : 1][KR];
  __shared__ float s_sc[kFields];
  uint2 *list = A.tlist2 + (size_t)blockIdx.x * A.tlist_cap;
  const int jx = morton_x(lane), jy = morton_y(lane);

  for(int q = tid; q < gwords; q += nt) grids[q] = 0;
  const unsigned long long t_wg0 = A.prof ? wall_clock64() : 0;
  while(true) {
    // the next trial from the launch's counter, longest first
    __syncthreads();  // the previous particle's use of the shared arrays is over
    if(tid == 0) s_item = atomicAdd(A.work_next, 1);
    __syncthreads();
    if(s_item >= A.nblk) {
      if(A.prof && tid == 0) {
        const unsigned long long tw = wall_clock64() - t_wg0;
        atomicAdd(&A.prof[13], tw);
        atomicMax(&A.prof[14], tw);
        atomicAdd(&A.prof[15], 1ull);
      }
      break;
    }
    const int bid = A.work_order[s_item];
    const int ti = A.trial_base + bid;
    const int j = A.slot_pos ? A.slot_pos[bid] : bid;
    const int p = A.tidx[ti];
    const int t = A.P.type[p];
    const size_t o0 = (size_t)j * (K + 1);
    unsigned long long tp0 = 0, tp1 = 0, tpa = 0, tpb = 0;
    if(A.prof && tid == 0) tp0 = wall_clock64();
    if(tid == 0) {
      int s = 0;
      for(int o = 0; o <= K; o++) {
        s_off[o] = s;
        s += max(A.log_n[o0 + o], 0);
      }
      s_off[K + 1] = s;
    }
    if(tid < kFields) {
      float mx = 0.f;
      for(int o = 0; o <= K; o++) mx = fmaxf(mx, A.log_tot[(o0 + o) * kFields + tid]);
      int e = 0;
      (void)frexpf(mx, &e);  // mx < 2^e
      const int kp = min(126, 30 - e);
      s_kp[tid] = kp;
      s_sc[tid] = ldexpf(1.0f, kp);
    }
    for(int q = tid; q < ntile; q += nt) tcnt[q] = 0;
    for(int q = tid; q < nmask; q += nt) tmask[q] = 0;
    __syncthreads();
    const int R = s_off[K + 1];
    // 1. the tile lists: counts, offsets, entries.  A thread takes T16_LU records at a time,
    // their block-table and quad loads issued together.
    auto records = [&](auto visit_record) {
      for(int q0 = tid; q0 < R; q0 += nt * T16_LU) {
        unsigned int rec[T16_LU];
        int ob[T16_LU], quad[T16_LU];
#pragma unroll
        for(int u = 0; u < T16_LU; u++) {
          const int q = q0 + u * nt;
          ob[u] = -1;
          rec[u] = 0;
          if(q < R) rec[u] = t16_record(A, s_off, K, o0, q, ob[u]);
        }
#pragma unroll
        for(int u = 0; u < T16_LU; u++) quad[u] = ob[u] >= 0 ? A.log_q[rec[u]] : 0;
#pragma unroll
        for(int u = 0; u < T16_LU; u++)
          if(ob[u] >= 0) visit_record(rec[u], ob[u], quad[u]);
      }
    };
    records([&](unsigned int rec, int o, int quad) {
      t16_visit(quad, L, [&](int m) {
        atomicAdd(&tcnt[m], 1);
        if(!(t16_mask_get<KR>(tmask, m) & (1u << o))) t16_mask_or<KR>(tmask, m, 1u << o);
      });
    });
    __syncthreads();
    if(wave == 0) t16_scan(tcnt, toff, tact, &s_nact, ntile, lane);
    __syncthreads();
    if(toff[ntile] > A.tlist_cap) {
      // cannot happen when the caller sizes tlist_cap from k_log_totals; reported, not evaluated
      for(int q = tid; q < ntile; q += nt) tcnt[q] = 0;
      if(tid == 0) {
        atomicOr(A.tile_err, 1u);
        A.accept[ti] = 0;
        A.dF[ti] = 0;
      }
      continue;
    }
    for(int q = tid; q < ntile; q += nt) tcnt[q] = 0;
    __syncthreads();
    records([&](unsigned int rec, int o, int quad) {
      t16_visit(quad, L, [&](int m) { list[toff[m] + atomicAdd(&tcnt[m], 1)] = make_uint2(rec, (unsigned int)o); });
    });
    __syncthreads();
    // the counts shared the grids' words: zero again for the tiles
    for(int q = tid; q < ntile; q += nt) tcnt[q] = 0;
    const int nact = s_nact;
    if(A.prof && tid == 0) {
      tp1 = wall_clock64();
      atomicAdd(&A.prof[0], tp1 - tp0);
      atomicAdd(&A.prof[5], (unsigned long long)nact);
      atomicAdd(&A.prof[6], (unsigned long long)toff[ntile]);
      atomicAdd(&A.prof[7], 1ull);
    }

    // 2. the tiles in Morton order
    const int f = wave;
    // (the pairs in FP32: the residual's rounding moves a term |b + D| - |b| only where b + D
    // changes sign, by at most 2^-23 of |D|)
    const float2 *bwf = A.merit_bwf + ((size_t)t * kFields + f) * Nstack;
    const int kp = s_kp[f];
    double acc[KR];
#pragma unroll
    for(int k = 0; k < KR; k++) acc[k] = 0.0;
    double accs = 0.0;  // lane k: trial k's terms of the tile nodes and above
    // lane 4q + j: the 4x4-node terms of trials j, j + 4, ...; lane 16r + k: trial k's 8x8 terms
    double acc2[(KR + 3) / 4], acc3 = 0.0;
#pragma unroll
    for(int u = 0; u < (KR + 3) / 4; u++) acc2[u] = 0.0;
    int *stk = &s_stk[f][0][0];  // [l][k], read and written by lane k of wave f only
    if(lane < KR)
      for(int l = 0; l < NSTK; l++) stk[l * KR + lane] = 0;
    float2 pcf[4] = {}, p1f = make_float2(0.f, 0.f), up = make_float2(0.f, 0.f);
    float2 p2f = p1f, p3f = p1f, p4f = p1f;  // the 4x4, 8x8 and tile nodes' pairs
    // (only the levels with cut-A nodes in the tile: the others' terms are not formed)
    auto load_pairs = [&](int m) {
      const int gz = t16_z(m) * T16 + 2 * jy, gr = t16_r(m) * T16 + 2 * jx;
      const unsigned int cmn = A.tcut[(size_t)t * ntile + m];
      if(cmn & 1u) {
        pcf[0] = bwf[GALIC_STACKOFFSET(L, gz, gr)];
        pcf[1] = bwf[GALIC_STACKOFFSET(L, gz, gr + 1)];
        pcf[2] = bwf[GALIC_STACKOFFSET(L, gz + 1, gr)];
        pcf[3] = bwf[GALIC_STACKOFFSET(L, gz + 1, gr + 1)];
      }
      if(cmn & 2u) p1f = bwf[GALIC_STACKOFFSET(L - 1, gz >> 1, gr >> 1)];
      if(cmn & 4u) p2f = bwf[GALIC_STACKOFFSET(L - 2, gz >> 2, gr >> 2)];
      if(cmn & 8u) p3f = bwf[GALIC_STACKOFFSET(L - 3, gz >> 3, gr >> 3)];
      if(cmn & 16u) p4f = bwf[GALIC_STACKOFFSET(L - 4, t16_z(m), t16_r(m))];
    };
    // closes trial k's open ancestors of the tile prev (all of them, or those m does not
    // share) from level lt - 1 upwards; up holds their pairs (lane l: level l)
    auto close_levels = [&](int prev, int m, bool all) {
      bool open = true;
#pragma unroll
      for(int l = NSTK - 1; l >= 0; l--) {
        if(l < lt && open) {
          const int sh = 2 * (lt - l);
          if(!all && (prev >> sh) == (m >> sh)) {
            open = false;
          } else {
            const double b = ldexp((double)__shfl(up.x, l), kp), w = (double)__shfl(up.y, l);
            const int sl = lane < KR ? stk[l * KR + lane] : 0;
            accs = t16_term(b, w, (double)sl, accs);
            if(lane < KR) {
              if(l > 0) stk[(l - 1) * KR + lane] += sl;
              stk[l * KR + lane] = 0;
            }
          }
        }
      }
    };
    // Phase A's list entries are fetched ahead: during tile ai's pyramid each thread loads its
    // first entry of tile ai+2, so a tile's adds start with the entry in registers.
    auto fetch_entry = [&](int ai2) {
      uint2 en = make_uint2(0u, 0xffffffffu);
      if(ai2 < nact) {
        const int m2 = tact[ai2], e = toff[m2] + tid;
        if(e < toff[m2 + 1]) en = list[e];
      }
      return en;
    };
    // the cells of one entry inside tile (tz, tr) into its log's grids
    auto add_values = [&](int o, int quad, const float4 (&x)[kFields], int tz, int tr) {
      const int lz = (quad >> L) - tz * T16, lr = (quad & (nb - 1)) - tr * T16;
#pragma unroll
      for(int c = 0; c < 4; c++) {
        const int cz = lz + (c >> 1), cr = lr + (c & 1);
        if((unsigned)cz < (unsigned)T16 && (unsigned)cr < (unsigned)T16) {
          const int w = t16_cell(cz, cr);
#pragma unroll
          for(int f2 = 0; f2 < kFields; f2++) {
            const float v = c == 0 ? x[f2].x : c == 1 ? x[f2].y : c == 2 ? x[f2].z : x[f2].w;
            if(v != 0.f) atomicAdd(&grids[(o * kFields + f2) * 256 + w], __float2int_rn(v * s_sc[f2]));
          }
        }
      }
    };
    auto add_entry = [&](uint2 en, int tz, int tr) {
      const unsigned int rec = en.x;
      const int blk = (int)(rec / LOG_B), jj = (int)(rec % LOG_B);
      float4 x[kFields];
#pragma unroll
      for(int f2 = 0; f2 < kFields; f2++) x[f2] = A.log_v[((size_t)blk * LOG_B + jj) * kFields + f2];
      add_values((int)en.y, A.log_q[rec], x, tz, tr);
    };
    uint2 pe_cur = fetch_entry(0);
    uint2 pe = fetch_entry(1);
    if(nact > 0) load_pairs(tact[0]);
    int prev = -1;
    for(int ai = 0; ai < nact; ai++) {
      const int m = tact[ai], tz = t16_z(m), tr = t16_r(m);
      if(A.prof && tid == 0) tpa = wall_clock64();
      // phase A: the tile's entries into the logs' grids (the first one's list entry fetched
      // ahead)
      if(pe_cur.y != 0xffffffffu) add_entry(pe_cur, tz, tr);
      for(int e = toff[m] + tid + nt; e < toff[m + 1]; e += nt) add_entry(list[e], tz, tr);
      __syncthreads();
      if(A.prof && tid == 0) {
        tpb = wall_clock64();
        atomicAdd(&A.prof[9], tpb - tpa);
      }
      // phase B: wave f, field f (the next tiles' entries on their way)
      pe_cur = pe;
      pe = fetch_entry(ai + 2);
      if(prev >= 0) close_levels(prev, m, false);
      if(lane < lt) {
        const int a = m >> (2 * (lt - lane));
        up = bwf[GALIC_STACKOFFSET(lane, t16_z(a), t16_r(a))];
      }
      {
        // the levels that hold cut-A nodes in this tile: only their terms can be nonzero (the
        // weights vanish off the cut), and the node sums are formed only if a level above the
        // cells needs them
        const unsigned int cm = A.tcut[(size_t)t * ntile + m];
        const bool cells = cm & 1u, node1 = cm & 2u, sums = cm & 0x3eu;
        // the 4x4, 8x8 and tile nodes' pairs of the levels on the cut (loaded a tile ahead with
        // the cells' and 2x2 nodes'), used after the trials
        const float2 p2 = p2f, p3 = p3f, p4 = p4f;
        // residuals in the particle's units
        double2 pc[4];
#pragma unroll
        for(int c = 0; c < 4; c++) pc[c] = make_double2(ldexp((double)pcf[c].x, kp), (double)pcf[c].y);
        const double2 p1 = make_double2(ldexp((double)p1f.x, kp), (double)p1f.y);
        const unsigned int tm = t16_mask_get<KR>(tmask, m);
        // The coarser nodes' sums of trial k are handed to one lane each and their terms formed
        // after the loop: the 4x4 node's (in all four lanes of its quad after the quad sum) to
        // lane 4q + k % 4, register k / 4; the 8x8 node's (in all 16 lanes of its row) to lane
        // 16r + k; the tile's to lane k.
        int s2x[(KR + 3) / 4], s3x = 0, tsk = 0;
#pragma unroll
        for(int u = 0; u < (KR + 3) / 4; u++) s2x[u] = 0;
        int4 *og = (int4 *)(grids + f * 256) + lane;  // O's word; trial k's: og + (k + 1) * kFields * 64
        const int4 o4 = *og;
        *og = make_int4(0, 0, 0, 0);
        int4 *g0 = og + kFields * 64;
        int4 gn = *g0;
#pragma unroll
        for(int k = 0; k < KR; k++) {
          if(k < K) {
            const int4 g4 = gn;
            if(k + 1 < K) gn = g0[(k + 1) * kFields * 64];
            // (a compiler barrier: trial k+1's word is read here, one trial ahead, not all
            // trials' words at once, which would hold 4 registers per trial)
            asm volatile("" ::: "memory");
            if(tm & (1u | (2u << k))) {
              g0[k * kFields * 64] = make_int4(0, 0, 0, 0);
              const int4 d = make_int4(g4.x - o4.x, g4.y - o4.y, g4.z - o4.z, g4.w - o4.w);
              double a = acc[k];
              if(cells) {
                a = t16_term(pc[0].x, pc[0].y, (double)d.x, a);
                a = t16_term(pc[1].x, pc[1].y, (double)d.y, a);
                a = t16_term(pc[2].x, pc[2].y, (double)d.z, a);
                a = t16_term(pc[3].x, pc[3].y, (double)d.w, a);
              }
              if(sums) {
                int s = (d.x + d.y) + (d.z + d.w);
                if(node1) a = t16_term(p1.x, p1.y, (double)s, a);
                s += dpp_i32<kDppQuadXor1>(s);
                s += dpp_i32<kDppQuadXor2>(s);
                s2x[k >> 2] = (lane & 3) == (k & 3) ? s : s2x[k >> 2];
                s += dpp_i32<kDppRowRor4>(s);
                s += dpp_i32<kDppRowRor8>(s);
                s3x = (lane & 15) == k ? s : s3x;
                const int s4 = __builtin_amdgcn_readlane(s, 0) + __builtin_amdgcn_readlane(s, 16) +
                               __builtin_amdgcn_readlane(s, 32) + __builtin_amdgcn_readlane(s, 48);
                tsk = lane == k ? s4 : tsk;
              }
              acc[k] = a;
            }
          }
        }
        // the coarser nodes' terms of the lane's trials (0 for trials beyond K or not evaluated)
        if(cm & 4u)
#pragma unroll
          for(int u = 0; u < (KR + 3) / 4; u++)
            acc2[u] = t16_term(ldexp((double)p2.x, kp), (double)p2.y, (double)s2x[u], acc2[u]);
        if(cm & 8u) acc3 = t16_term(ldexp((double)p3.x, kp), (double)p3.y, (double)s3x, acc3);
        if(cm & 16u) accs = t16_term(ldexp((double)p4.x, kp), (double)p4.y, (double)tsk, accs);
        if((cm & 32u) && lt > 0 && lane < KR) stk[(lt - 1) * KR + lane] += tsk;
      }
      prev = m;
      if(ai + 1 < nact) load_pairs(tact[ai + 1]);
      __syncthreads();
      if(A.prof && tid == 0) atomicAdd(&A.prof[10], wall_clock64() - tpb);
    }
    if(prev >= 0) close_levels(prev, prev, true);
    // per-trial totals of field f
    {
      const double MT = A.MType[t];
      const double coef = f == 0 ? 1.0 / MT : (f == 1 ? A.fac[t][0] : f == 2 ? A.fac[t][1] : f == 3 ? A.fac[t][2]
                                                                                                  : A.fac[t][3]) /
                                                  MT / 4;
      const double unit = ldexp(coef, -kp);
#pragma unroll
      for(int k = 0; k < KR; k++)
        if(k < K) {
          double v = acc[k];
          if((lane & 3) == (k & 3)) v += acc2[k >> 2];
          if((lane & 15) == k) v += acc3;
          if(lane == k) v += accs;
          v = wave_sum(v);
          if(lane == 0) s_part[f][k] = v * unit;
        }
    }
    __syncthreads();
    // 3. the choice (the taken trial's deltas are the region commit's, k_commit_*)
    for(int k = tid; k < K; k += nt) {
      const size_t q = (size_t)ti * K + k;
      const double *fit = A.fit + (size_t)t * 16;
      double orbit_terms = 0;
#pragma unroll
      for(int f2 = 0; f2 < kFields; f2++) orbit_terms += s_part[f2][k];
      s_dF[k] = orbit_terms + fit[5] / fit[6] * A.vinst[q];
      A.dFk[q] = s_dF[k];
    }
    __syncthreads();
    if(tid == 0) {
      int bk = -1;
      double bv = 0;
      for(int k = 0; k < K; k++)
        if(s_dF[k] < bv) {
          bv = s_dF[k];
          bk = k;
        }
      s_best = bk;
      A.accept[ti] = bk + 1;
      A.dF[ti] = (bk >= 0) ? bv : (K > 0 ? s_dF[0] : 0.0);
      if(A.prof) {
        const unsigned long long te = wall_clock64();
        atomicAdd(&A.prof[1], te - tp1);
        if(bk >= 0) atomicAdd(&A.prof[8], 1ull);
        atomicMax(&A.prof[4], te - tp0);
        atomicAdd(&A.prof[11], (te - tp0) * (te - tp0) / 100);
        atomicMax(&A.prof[12], (unsigned long long)toff[ntile]);
      }
    }
  }
}

template __global__ void k_eval_t16<0, 8>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_t16<9, 8>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_t16<0, 16>(DevModel M, EvalArgs32 A);
template __global__ void k_eval_t16<9, 16>(DevModel M, EvalArgs32 A);

// ---------------------------------------------------- fixed point commit --

// Loc += delta in fixed point; the FP64 grid the global fit reads is Loc/scale.
__global__ void k_commit_q(long long *locq, long long *dq, double *loc, size_t n_per_field, int nfield,
                           const double *inv_scale, size_t n) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= n) return;
  int f = (int)((k / n_per_field) % nfield);
  long long v = locq[k] + dq[k];
  locq[k] = v;
  dq[k] = 0;
  loc[k] = (double)v * inv_scale[f];
}

// Instantaneous moments of every particle in fixed point (main.c:226-229).
__global__ void k_inst_init_q(DevModel M, Particles P, long long *locIq, long long *locMq, int ngrid, double iscale,
                              double mscale) {
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
  long long *li = locIq + (size_t)t * 4 * ngrid;
  for(int k = 0; k < 4; k++) {
    atomic_q(&locMq[(size_t)t * ngrid + c[k]], mscale * w[k] * mass);
    for(int f = 0; f < 4; f++) atomic_q(&li[(size_t)f * ngrid + c[k]], iscale * w[k] * val[f]);
  }
}

}  // namespace galic
