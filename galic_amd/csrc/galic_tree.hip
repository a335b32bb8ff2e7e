// galic_tree.hip -- Barnes-Hut fill of the force and potential tables on gfx950.
//
// The reference builds an oct-tree by inserting particles one at a time
// (forcetree.c:369-495), sums monopole moments recursively
// (forcetree.c:591-722), and walks it once per target with a geometric opening
// criterion and Gadget-2 spline softening (forcetree_walk.c:12-222).  The tree
// that insertion produces is unique: a node exists for every cube of the
// subdivision that holds at least two particles, down to the depth where they
// separate.  Here the same tree is built in parallel:
//   1. integer positions exactly as forcetree.c:199-201 (52 mantissa bits of
//      (x - corner)/len + 1), Morton keys from their top 21 bits per axis with
//      the child numbering x + 2y + 4z of forcetree.c:216-217;
//   2. radix sort of the keys;
//   3. level by level, every node splits its sorted particle range into the 8
//      child ranges by binary search; ranges of >= 2 particles become nodes;
//   4. moments bottom-up in FP64, in child order 0..7 as the reference;
//   5. nextnode/sibling threading top-down, giving the reference's depth-first
//      walk order;
//   6. one target per lane walks the tree without a stack.
// Particles that share all 21 key levels (closer than len/2^21) stay together
// in a leaf bucket; opening it interacts with them one by one.  The reference
// would keep splitting; this changes results only for targets within that
// distance of such a pair.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/galic_hip.h"
#include "galic_ctx.h"
#include "galic_kernels.h"

namespace galic {
galic_ctx *bound_ctx();
Ctx &ctx_state(galic_ctx *ctx);
namespace tree {

constexpr int kMaxDepth = 21;

struct Tree {
  int nsrc = 0;
  int cap = 0;
  // sorted sources
  double *x = nullptr, *y = nullptr, *z = nullptr, *m = nullptr;
  unsigned long long *key = nullptr;
  int *nextp = nullptr;  // next walk entry after each particle
  // nodes
  int *first = nullptr, *count = nullptr, *depth = nullptr;
  int *child = nullptr;  // [node][8]: -1, particle (< nsrc) or nsrc + node
  double *mass = nullptr, *sx = nullptr, *sy = nullptr, *sz = nullptr;
  double *quad = nullptr;  // [6][node] traceless quadrupole xx, yy, zz, xy, xz, yz (with quadrupole moments)
  float4 *quadf = nullptr; // [node][2] the same in FP32 for the FP32 walk: (xx, yy, zz, xy), (xz, yz, 0, 0)
  float *len = nullptr;
  int *sibling = nullptr, *nextnode = nullptr, *bucket = nullptr;
  int *counter = nullptr;
  std::vector<int> level_lo;  // nodes of depth d are [level_lo[d], level_lo[d+1])
};

__device__ __forceinline__ unsigned long long spread3(unsigned long long v) {
  v &= 0x1fffffULL;
  v = (v | (v << 32)) & 0x1f00000000ffffULL;
  v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}

__device__ __forceinline__ unsigned long long mantissa(double d) {
  return (unsigned long long)__double_as_longlong(d) & 0xFFFFFFFFFFFFFull;
}

// forcetree.c:197-210 integer positions; Morton key from bits 51..31 of each
__global__ void k_keys(int n, const double *px, const double *py, const double *pz, double cx, double cy, double cz,
                       double inv_len, unsigned long long *key, int *idx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  unsigned long long xb = mantissa(((px[i] - cx) * inv_len) + 1.0) >> 31;
  unsigned long long yb = mantissa(((py[i] - cy) * inv_len) + 1.0) >> 31;
  unsigned long long zb = mantissa(((pz[i] - cz) * inv_len) + 1.0) >> 31;
  key[i] = spread3(xb) | (spread3(yb) << 1) | (spread3(zb) << 2);
  idx[i] = i;
}

__global__ void k_gather(int n, const int *idx, const double *px, const double *py, const double *pz,
                         const double *pm, double *x, double *y, double *z, double *m) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  int j = idx[i];
  x[i] = px[j];
  y[i] = py[j];
  z[i] = pz[j];
  m[i] = pm[j];
}

__device__ __forceinline__ int digit_at(unsigned long long key, int d) {  // child digit below depth d
  return (int)((key >> (3 * (kMaxDepth - 1 - d))) & 7ull);
}

// Split every node of depth d into its 8 child ranges (binary search on the
// sorted keys); ranges of >= 2 particles become nodes of depth d+1.
__global__ void k_split(int lo, int hi, int d, int nsrc, int cap, const unsigned long long *key, int *first,
                        int *count, int *depth, int *child, int *bucket, int *counter) {
  int nd = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if(nd >= hi) return;
  int f = first[nd], c = count[nd];
  int *ch = child + (size_t)nd * 8;
  for(int k = 0; k < 8; k++) ch[k] = -1;
  if(c >= 2 && d >= kMaxDepth) {  // particles closer than the key resolution: a leaf bucket
    bucket[nd] = 1;
    return;
  }
  bucket[nd] = 0;
  int b[9];
  b[0] = f;
  b[8] = f + c;
  for(int dg = 1; dg < 8; dg++) {
    int l = b[dg - 1], r = f + c;  // first index in [l, r) with digit >= dg
    while(l < r) {
      int mid = (l + r) >> 1;
      if(digit_at(key[mid], d) < dg)
        l = mid + 1;
      else
        r = mid;
    }
    b[dg] = l;
  }
  for(int dg = 0; dg < 8; dg++) {
    int cnt = b[dg + 1] - b[dg];
    if(cnt == 1) {
      ch[dg] = b[dg];
    } else if(cnt >= 2) {
      int id = atomicAdd(counter, 1);
      if(id >= cap) {
        ch[dg] = -2;  // overflow, reported by the host
        continue;
      }
      first[id] = b[dg];
      count[id] = cnt;
      depth[id] = d + 1;
      ch[dg] = nsrc + id;
    }
  }
}

// Traceless quadrupole m (3 d_i d_j - d^2 delta_ij) of a point mass at offset d,
// added to q (xx, yy, zz, xy, xz, yz).
__device__ __forceinline__ void add_quad(double q[6], double m, double dx, double dy, double dz) {
  const double d2 = dx * dx + dy * dy + dz * dz;
  q[0] += m * (3 * dx * dx - d2);
  q[1] += m * (3 * dy * dy - d2);
  q[2] += m * (3 * dz * dz - d2);
  q[3] += m * 3 * dx * dy;
  q[4] += m * 3 * dx * dz;
  q[5] += m * 3 * dy * dz;
}

// Monopole moments of the nodes of one depth (forcetree.c:619-697), children in order 0..7.
// With quad (nnodes per component), also the traceless quadrupole about the
// centre of mass, from the children's own quadrupoles shifted to it.
__global__ void k_moments(int lo, int hi, int nsrc, int nnodes, const int *child, const int *bucket, const int *first,
                          const int *count, const double *x, const double *y, const double *z, const double *m,
                          double *mass, double *sx, double *sy, double *sz, double *quad) {
  int nd = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if(nd >= hi) return;
  double ms = 0, s0 = 0, s1 = 0, s2 = 0;
  if(bucket[nd]) {
    for(int p = first[nd]; p < first[nd] + count[nd]; p++) {
      ms += m[p];
      s0 += m[p] * x[p];
      s1 += m[p] * y[p];
      s2 += m[p] * z[p];
    }
  } else {
    const int *ch = child + (size_t)nd * 8;
    for(int k = 0; k < 8; k++) {
      int c = ch[k];
      if(c < 0) continue;
      if(c < nsrc) {
        ms += m[c];
        s0 += m[c] * x[c];
        s1 += m[c] * y[c];
        s2 += m[c] * z[c];
      } else {
        int q = c - nsrc;
        ms += mass[q];
        s0 += mass[q] * sx[q];
        s1 += mass[q] * sy[q];
        s2 += mass[q] * sz[q];
      }
    }
  }
  if(ms) {
    s0 /= ms;
    s1 /= ms;
    s2 /= ms;
  }
  mass[nd] = ms;
  sx[nd] = s0;
  sy[nd] = s1;
  sz[nd] = s2;
  if(!quad) return;
  double q[6] = {0, 0, 0, 0, 0, 0};
  if(bucket[nd]) {
    for(int p = first[nd]; p < first[nd] + count[nd]; p++) add_quad(q, m[p], x[p] - s0, y[p] - s1, z[p] - s2);
  } else {
    const int *ch = child + (size_t)nd * 8;
    for(int k = 0; k < 8; k++) {
      int c = ch[k];
      if(c < 0) continue;
      if(c < nsrc) {
        add_quad(q, m[c], x[c] - s0, y[c] - s1, z[c] - s2);
      } else {
        int r = c - nsrc;
        for(int j = 0; j < 6; j++) q[j] += quad[(size_t)j * nnodes + r];
        add_quad(q, mass[r], sx[r] - s0, sy[r] - s1, sz[r] - s2);
      }
    }
  }
  for(int j = 0; j < 6; j++) quad[(size_t)j * nnodes + nd] = q[j];
}

// nextnode = first child; the sibling of a child is the next child, the last
// child inherits its parent's sibling (forcetree.c:624-638).
__global__ void k_thread(int lo, int hi, int nsrc, const int *child, const int *bucket, int *sibling, int *nextnode,
                         int *nextp) {
  int nd = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if(nd >= hi) return;
  int sib = sibling[nd];
  if(bucket[nd]) {
    nextnode[nd] = sib;  // opened buckets are walked inline
    return;
  }
  const int *ch = child + (size_t)nd * 8;
  int list[8], n = 0;
  for(int k = 0; k < 8; k++)
    if(ch[k] >= 0) list[n++] = ch[k];
  nextnode[nd] = n ? list[0] : sib;
  for(int k = 0; k < n; k++) {
    int nx = (k + 1 < n) ? list[k + 1] : sib;
    int c = list[k];
    if(c < nsrc)
      nextp[c] = nx;
    else
      sibling[c - nsrc] = nx;
  }
}

// FP32 copy of the quadrupoles, two 16-byte loads per node in the FP32 walk
__global__ void k_quadf(int nnodes, const double *quad, float4 *quadf) {
  int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if(nd >= nnodes) return;
  const size_t nn = nnodes;
  quadf[2 * nd] = make_float4((float)quad[nd], (float)quad[nn + nd], (float)quad[2 * nn + nd], (float)quad[3 * nn + nd]);
  quadf[2 * nd + 1] = make_float4((float)quad[4 * nn + nd], (float)quad[5 * nn + nd], 0.f, 0.f);
}

__global__ void k_len(int lo, int hi, double dlen, float *len) {
  int nd = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if(nd < hi) len[nd] = (float)dlen;
}

struct WalkArgs {
  int nsrc, ntarget;
  const double *tx, *ty, *tz;
  const double *x, *y, *z, *m;
  const int *nextp, *first, *count, *bucket, *sibling, *nextnode;
  const double *mass, *sx, *sy, *sz;
  const double *quad;   // [6][nnodes] or null: monopole only
  const float4 *quadf;  // [nnodes][2] FP32 copy (FP32 walk)
  int nnodes;
  const float *len;
  const int *child;     // [node][8] (the group walk)
  int blk_nb;           // group walk: targets are the (section, z, R) table nodes of nb x nb grids,
                        // a group an 8 x 8 block of one section; 0: 64 consecutive targets
  double theta, h, G;
  int plummer;          // Plummer softening of length eps instead of the spline of length h
  double eps;
  double *ax, *ay, *az, *pot;
  unsigned long long *nint;
};

// Gadget-2 cubic spline monopole (forcetree_walk.c:158-191)
__device__ __forceinline__ void interact64(double mass, double r2, double h, double &fac, double &wp) {
  double r = sqrt(r2);
  if(r >= h) {
    fac = mass / (r2 * r);
    wp = -mass / r;
  } else {
    double h_inv = 1.0 / h;
    double h3_inv = h_inv * h_inv * h_inv;
    double u = r * h_inv;
    if(u < 0.5) {
      fac = mass * h3_inv * (10.666666666667 + u * u * (32.0 * u - 38.4));
      wp = mass * h_inv * (-2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6)));
    } else {
      fac = mass * h3_inv *
            (21.333333333333 - 48.0 * u + 38.4 * u * u - 10.666666666667 * u * u * u - 0.066666666667 / (u * u * u));
      wp = mass * h_inv *
           (-3.2 + 0.066666666667 / u + u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u))));
    }
  }
}

__device__ __forceinline__ void interact32(float mass, float r2, float h, float &fac, float &wp) {
  float rinv = (r2 > 0.f) ? __builtin_amdgcn_rsqf(r2) : 0.f;
  float r = r2 * rinv;  // 0 for coincident points, which then take the u = 0 branch
  if(r >= h) {
    fac = mass * rinv * rinv * rinv;
    wp = -mass * rinv;
  } else {
    float h_inv = 1.0f / h;
    float h3_inv = h_inv * h_inv * h_inv;
    float u = r * h_inv;
    if(u < 0.5f) {
      fac = mass * h3_inv * (10.666666666667f + u * u * (32.0f * u - 38.4f));
      wp = mass * h_inv * (-2.8f + u * u * (5.333333333333f + u * u * (6.4f * u - 9.6f)));
    } else {
      float ui = 1.0f / u;
      fac = mass * h3_inv *
            (21.333333333333f - 48.0f * u + 38.4f * u * u - 10.666666666667f * u * u * u - 0.066666666667f * ui * ui * ui);
      wp = mass * h_inv *
           (-3.2f + 0.066666666667f * ui + u * u * (10.666666666667f + u * (-16.0f + u * (9.6f - 2.133333333333f * u))));
    }
  }
}

// Plummer-softened monopole: potential -m / sqrt(r^2 + eps^2)
__device__ __forceinline__ void plummer64(double mass, double r2, double eps, double &fac, double &wp) {
  const double s2 = r2 + eps * eps;
  const double sinv = 1.0 / sqrt(s2);
  fac = mass * sinv * sinv * sinv;
  wp = -mass * sinv;
}

__device__ __forceinline__ void plummer32(float mass, float r2, float eps, float &fac, float &wp) {
  const float sinv = __builtin_amdgcn_rsqf(r2 + eps * eps);
  fac = mass * sinv * sinv * sinv;
  wp = -mass * sinv;
}

// Quadrupole term of a node at offset d = (centre of mass - target), distance^2
// s2 (softened for Plummer): potential -(1/2) d.Q.d / s^5, acceleration
// -Q.d / s^5 + (5/2) (d.Q.d) d / s^7.
template <typename T>
__device__ __forceinline__ void quad_term(const T q[6], T dx, T dy, T dz, T s2, T sinv, T &ax, T &ay, T &az, T &pot) {
  const T qx = q[0] * dx + q[3] * dy + q[4] * dz;
  const T qy = q[3] * dx + q[1] * dy + q[5] * dz;
  const T qz = q[4] * dx + q[5] * dy + q[2] * dz;
  const T dqd = dx * qx + dy * qy + dz * qz;
  const T s5 = sinv * sinv * sinv * sinv * sinv;
  const T c = (T)2.5 * dqd * s5 / s2;
  ax += -qx * s5 + c * dx;
  ay += -qy * s5 + c * dy;
  az += -qz * s5 + c * dz;
  pot += (T)-0.5 * dqd * s5;
}

// forcetree_walk.c:58-192 for one target per lane (mode 0, one task)
__global__ void __launch_bounds__(256) k_walk64(WalkArgs A) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= A.ntarget) return;
  const double px = A.tx[t], py = A.ty[t], pz = A.tz[t];
  double acc_x = 0, acc_y = 0, acc_z = 0, potv = 0, fac, wp;
  unsigned long long nint = 0;
  int no = A.nsrc;  // root
  while(no >= 0) {
    double dx, dy, dz, r2, mass;
    if(no < A.nsrc) {
      dx = A.x[no] - px;
      dy = A.y[no] - py;
      dz = A.z[no] - pz;
      r2 = dx * dx + dy * dy + dz * dz;
      mass = A.m[no];
      no = A.nextp[no];
    } else {
      int q = no - A.nsrc;
      mass = A.mass[q];
      dx = A.sx[q] - px;
      dy = A.sy[q] - py;
      dz = A.sz[q] - pz;
      r2 = dx * dx + dy * dy + dz * dz;
      float l = A.len[q];
      float l2 = l * l;
      if(l2 > r2 * A.theta * A.theta) {
        if(A.bucket[q]) {
          for(int p = A.first[q]; p < A.first[q] + A.count[q]; p++) {
            double ex = A.x[p] - px, ey = A.y[p] - py, ez = A.z[p] - pz;
            double e2 = ex * ex + ey * ey + ez * ez;
            if(A.m[p]) {
              if(A.plummer)
                plummer64(A.m[p], e2, A.eps, fac, wp);
              else
                interact64(A.m[p], e2, A.h, fac, wp);
              acc_x += ex * fac;
              acc_y += ey * fac;
              acc_z += ez * fac;
              potv += wp;
              nint++;
            }
          }
          no = A.sibling[q];
        } else {
          no = A.nextnode[q];
        }
        continue;
      }
      no = A.sibling[q];
      if(A.quad && mass) {
        const size_t nn = A.nnodes;
        const double qq[6] = {A.quad[q], A.quad[nn + q], A.quad[2 * nn + q],
                              A.quad[3 * nn + q], A.quad[4 * nn + q], A.quad[5 * nn + q]};
        const double s2 = A.plummer ? r2 + A.eps * A.eps : r2;
        if(A.plummer || s2 >= A.h * A.h) quad_term(qq, dx, dy, dz, s2, 1.0 / sqrt(s2), acc_x, acc_y, acc_z, potv);
      }
    }
    if(mass) {
      if(A.plummer)
        plummer64(mass, r2, A.eps, fac, wp);
      else
        interact64(mass, r2, A.h, fac, wp);
      acc_x += dx * fac;
      acc_y += dy * fac;
      acc_z += dz * fac;
      potv += wp;
      nint++;
    }
  }
  A.ax[t] = acc_x * A.G;
  A.ay[t] = acc_y * A.G;
  A.az[t] = acc_z * A.G;
  A.pot[t] = potv * A.G;
  atomicAdd(A.nint, nint);
}

// The same walk with FP32 interactions: separations are formed in FP64 (node
// centres of mass and particles stay FP64) and rounded once, the kernel and the
// sums run in FP32 per target, totals in FP64.
__global__ void __launch_bounds__(256) k_walk32(WalkArgs A) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= A.ntarget) return;
  const double px = A.tx[t], py = A.ty[t], pz = A.tz[t];
  float acc_x = 0, acc_y = 0, acc_z = 0, potv = 0, fac, wp;
  const float h = (float)A.h, th2 = (float)(A.theta * A.theta), eps = (float)A.eps;
  unsigned int nint = 0;
  int no = A.nsrc;
  while(no >= 0) {
    float dx, dy, dz, r2, mass;
    if(no < A.nsrc) {
      dx = (float)(A.x[no] - px);
      dy = (float)(A.y[no] - py);
      dz = (float)(A.z[no] - pz);
      r2 = dx * dx + dy * dy + dz * dz;
      mass = (float)A.m[no];
      no = A.nextp[no];
    } else {
      int q = no - A.nsrc;
      mass = (float)A.mass[q];
      dx = (float)(A.sx[q] - px);
      dy = (float)(A.sy[q] - py);
      dz = (float)(A.sz[q] - pz);
      r2 = dx * dx + dy * dy + dz * dz;
      float l = A.len[q];
      if(l * l > r2 * th2) {
        if(A.bucket[q]) {
          for(int p = A.first[q]; p < A.first[q] + A.count[q]; p++) {
            float ex = (float)(A.x[p] - px), ey = (float)(A.y[p] - py), ez = (float)(A.z[p] - pz);
            float e2 = ex * ex + ey * ey + ez * ez;
            if(A.m[p] != 0) {
              if(A.plummer)
                plummer32((float)A.m[p], e2, eps, fac, wp);
              else
                interact32((float)A.m[p], e2, h, fac, wp);
              acc_x += ex * fac;
              acc_y += ey * fac;
              acc_z += ez * fac;
              potv += wp;
              nint++;
            }
          }
          no = A.sibling[q];
        } else {
          no = A.nextnode[q];
        }
        continue;
      }
      no = A.sibling[q];
      if(A.quadf && mass != 0) {
        const float4 qa = A.quadf[2 * q], qb = A.quadf[2 * q + 1];
        const float qq[6] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y};
        const float s2 = A.plummer ? r2 + eps * eps : r2;
        if(A.plummer || s2 >= h * h)
          quad_term(qq, dx, dy, dz, s2, __builtin_amdgcn_rsqf(s2), acc_x, acc_y, acc_z, potv);
      }
    }
    if(mass != 0) {
      if(A.plummer)
        plummer32(mass, r2, eps, fac, wp);
      else
        interact32(mass, r2, h, fac, wp);
      acc_x += dx * fac;
      acc_y += dy * fac;
      acc_z += dz * fac;
      potv += wp;
      nint++;
    }
  }
  A.ax[t] = (double)acc_x * A.G;
  A.ay[t] = (double)acc_y * A.G;
  A.az[t] = (double)acc_z * A.G;
  A.pot[t] = (double)potv * A.G;
  atomicAdd(A.nint, (unsigned long long)nint);
}

// Table projection of grid.c:719-750: radial and vertical components summed
// over the FG_SECTIONS azimuthal sections in order, then averaged.
__global__ void k_project(int nb, const double *ax, const double *ay, const double *az, const double *pt,
                          const double *cs, const double *sn, double *dpotdr, double *dpotdz, double *pot) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= nb * nb) return;
  double r = 0, zz = 0, p = 0;
  for(int s = 0; s < GALIC_FG_SECTIONS; s++) {
    size_t n = (size_t)s * nb * nb + i;
    r += ax[n] * cs[s] + ay[n] * sn[s];
    zz += az[n];
    p += pt[n];
  }
  dpotdr[i] = r / GALIC_FG_SECTIONS;
  dpotdz[i] = zz / GALIC_FG_SECTIONS;
  pot[i] = p / GALIC_FG_SECTIONS;
}

// Table-fill targets (grid.c:680-704): node (s, k, j) at (R_j cos phi_s, R_j sin phi_s, z_k)
// with R_j = z_j = rj[j], the same products as the host-built targets.
__global__ void k_fill_targets(int nb, const double *rj, const double *cs, const double *sn, double *tx, double *ty,
                               double *tz) {
  const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(n >= (size_t)GALIC_FG_SECTIONS * nb * nb) return;
  const int s = (int)(n / ((size_t)nb * nb)), k = (int)((n / nb) % nb), j = (int)(n % nb);
  tx[n] = rj[j] * cs[s];
  ty[n] = rj[j] * sn[s];
  tz[n] = rj[k];
}

// The FP32 orbit kernels' force table (DPotDR, DPotDz) from the FP64 tables.
__global__ void k_acc_f(int n, const double *dr, const double *dz, float2 *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) out[i] = make_float2((float)dr[i], (float)dz[i]);
}

// Direct summation in FP64 with the walk's interaction (the check of the
// tree's force error): one target per lane, sources staged through LDS.
__global__ void __launch_bounds__(256) k_direct64(int nsrc, const double *sx, const double *sy, const double *sz,
                                                   const double *sm, int ntarget, const double *tx, const double *ty,
                                                   const double *tz, double h, int plummer, double eps, double G,
                                                   double *ax, double *ay, double *az, double *pot) {
  __shared__ double4 tile[256];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < ntarget;
  const double px = live ? tx[t] : 0, py = live ? ty[t] : 0, pz = live ? tz[t] : 0;
  double acc_x = 0, acc_y = 0, acc_z = 0, potv = 0, fac, wp;
  for(int base = 0; base < nsrc; base += 256) {
    const int q = base + threadIdx.x;
    tile[threadIdx.x] = (q < nsrc) ? make_double4(sx[q], sy[q], sz[q], sm[q]) : make_double4(0, 0, 0, 0);
    __syncthreads();
    const int nt = min(256, nsrc - base);
    for(int k = 0; k < nt; k++) {
      const double4 s = tile[k];
      if(s.w != 0) {
        const double dx = s.x - px, dy = s.y - py, dz = s.z - pz;
        if(plummer)
          plummer64(s.w, dx * dx + dy * dy + dz * dz, eps, fac, wp);
        else
          interact64(s.w, dx * dx + dy * dy + dz * dz, h, fac, wp);
        acc_x += dx * fac;
        acc_y += dy * fac;
        acc_z += dz * fac;
        potv += wp;
      }
    }
    __syncthreads();
  }
  if(live) {
    ax[t] = acc_x * G;
    ay[t] = acc_y * G;
    az[t] = acc_z * G;
    pot[t] = potv * G;
  }
}

}  // namespace tree
}  // namespace galic

namespace {

using namespace galic::tree;

int terr(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return galic::set_error(code, buf);
}

#define TCHK(expr)                                                                                 \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if(e_ != hipSuccess) {                                                                         \
      rc = terr(e_ == hipErrorOutOfMemory ? GALIC_ERR_NOMEM : GALIC_ERR_HIP, "%s: %s", #expr,     \
                hipGetErrorString(e_));                                                            \
      goto done;                                                                                   \
    }                                                                                              \
  } while(0)

inline int nb256(size_t n) { return (int)((n + 255) / 256); }

// Device times of the last tree call of this process (ms): build (sort, split,
// moments, threading), walk, and the interaction count.
double g_last_build_ms = 0, g_last_walk_ms = 0, g_last_interactions = 0;

struct DevBufs {
  std::vector<void *> ptrs;
  template <typename T>
  hipError_t alloc(T **p, size_t n) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
    if(e == hipSuccess) ptrs.push_back((void *)*p);
    return e;
  }
  ~DevBufs() {
    for(void *p : ptrs) (void)hipFree(p);
  }
};

// Moments bottom-up (monopole, and the quadrupole where T.quad is allocated), node side
// lengths DomainLen / 2^depth (NODE.len, allvars.h:1095, held in float), then the
// nextnode/sibling threading top-down (forcetree.c:591-722).
int tree_moments_dev(hipStream_t st, Tree &T, double dlen) {
  const int nnodes = T.level_lo.back(), nsrc = T.nsrc;
  const int ndepth = (int)T.level_lo.size() - 1;
  for(int d = ndepth - 1; d >= 0; d--) {
    int lo = T.level_lo[d], hi = T.level_lo[d + 1];
    if(hi <= lo) continue;
    hipLaunchKernelGGL(k_moments, dim3(nb256(hi - lo)), dim3(256), 0, st, lo, hi, nsrc, nnodes,
                       (const int *)T.child, (const int *)T.bucket, (const int *)T.first, (const int *)T.count,
                       (const double *)T.x, (const double *)T.y, (const double *)T.z, (const double *)T.m, T.mass,
                       T.sx, T.sy, T.sz, T.quad);
    hipLaunchKernelGGL(k_len, dim3(nb256(hi - lo)), dim3(256), 0, st, lo, hi, ldexp(dlen, -d), T.len);
    if(hipGetLastError() != hipSuccess) return terr(GALIC_ERR_HIP, "tree moments launch failed");
  }
  int minus1 = -1;
  if(hipMemcpyAsync(T.sibling, &minus1, sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess)
    return terr(GALIC_ERR_HIP, "tree threading: copy failed");
  for(int d = 0; d < ndepth; d++) {
    int lo = T.level_lo[d], hi = T.level_lo[d + 1];
    if(hi <= lo) continue;
    hipLaunchKernelGGL(k_thread, dim3(nb256(hi - lo)), dim3(256), 0, st, lo, hi, nsrc, (const int *)T.child,
                       (const int *)T.bucket, T.sibling, T.nextnode, T.nextp);
    if(hipGetLastError() != hipSuccess) return terr(GALIC_ERR_HIP, "tree threading launch failed");
  }
  if(T.quad && T.quadf) {
    hipLaunchKernelGGL(k_quadf, dim3(nb256(nnodes)), dim3(256), 0, st, nnodes, (const double *)T.quad, T.quadf);
    if(hipGetLastError() != hipSuccess) return terr(GALIC_ERR_HIP, "tree quadrupole copy failed");
  }
  return GALIC_OK;
}

// Builds the tree of nsrc sources (device SoA inputs, domain cube at corner with side dlen)
// into T, its buffers owned by B.  sorted (optional) receives the device array giving, for each
// sorted source, its index among the inputs.
int tree_build_dev(const galic_params *p, hipStream_t st, int nsrc, const double *dpx, const double *dpy,
                   const double *dpz, const double *dpm, const double corner[3], double dlen, int precision,
                   DevBufs &B, Tree &T, int **sorted, float *build_ms) {
  int rc = GALIC_OK;
  unsigned long long *keys_in = nullptr;
  int *idx_in = nullptr, *idx = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  hipEvent_t eb = nullptr, ee = nullptr;
  (void)p;
  T.nsrc = nsrc;
  TCHK(hipEventCreate(&eb));
  TCHK(hipEventCreate(&ee));
  TCHK(hipEventRecord(eb, st));
  TCHK(B.alloc(&keys_in, nsrc));
  TCHK(B.alloc(&T.key, nsrc));
  TCHK(B.alloc(&idx_in, nsrc));
  TCHK(B.alloc(&idx, nsrc));
  TCHK(B.alloc(&T.x, nsrc));
  TCHK(B.alloc(&T.y, nsrc));
  TCHK(B.alloc(&T.z, nsrc));
  TCHK(B.alloc(&T.m, nsrc));
  TCHK(B.alloc(&T.nextp, nsrc));
  if(nsrc > 0) {
    hipLaunchKernelGGL(k_keys, dim3(nb256(nsrc)), dim3(256), 0, st, nsrc, dpx, dpy, dpz, corner[0], corner[1],
                       corner[2], 1.0 / dlen, keys_in, idx_in);
    TCHK(hipGetLastError());
    TCHK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys_in, T.key, idx_in, idx, (unsigned int)nsrc, 0, 63, st));
    TCHK(B.alloc((char **)&tmp, tmp_bytes));
    TCHK(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, T.key, idx_in, idx, (unsigned int)nsrc, 0, 63, st));
    hipLaunchKernelGGL(k_gather, dim3(nb256(nsrc)), dim3(256), 0, st, nsrc, (const int *)idx, dpx, dpy, dpz, dpm, T.x,
                       T.y, T.z, T.m);
    TCHK(hipGetLastError());
  }
  if(sorted) *sorted = idx;
  {
    // nodes: the unique insertion tree has fewer than 2 nsrc nodes unless
    // particles pair up at great depth; grow and rebuild if needed
    // (the reference raises TreeAllocFactor and retries, forcetree.c:48-70)
    T.cap = 2 * nsrc + 64;
    for(int attempt = 0; attempt < 6; attempt++) {
      TCHK(B.alloc(&T.first, T.cap));
      TCHK(B.alloc(&T.count, T.cap));
      TCHK(B.alloc(&T.depth, T.cap));
      TCHK(B.alloc(&T.child, (size_t)T.cap * 8));
      TCHK(B.alloc(&T.bucket, T.cap));
      TCHK(B.alloc(&T.counter, 1));
      int root[3] = {0, nsrc, 0}, one = 1;
      TCHK(hipMemcpyAsync(T.first, &root[0], sizeof(int), hipMemcpyHostToDevice, st));
      TCHK(hipMemcpyAsync(T.count, &root[1], sizeof(int), hipMemcpyHostToDevice, st));
      TCHK(hipMemcpyAsync(T.depth, &root[2], sizeof(int), hipMemcpyHostToDevice, st));
      TCHK(hipMemcpyAsync(T.counter, &one, sizeof(int), hipMemcpyHostToDevice, st));
      T.level_lo.assign(1, 0);
      int lo = 0, hi = 1, d = 0;
      bool overflow = false;
      while(lo < hi) {
        T.level_lo.push_back(hi);
        hipLaunchKernelGGL(k_split, dim3(nb256(hi - lo)), dim3(256), 0, st, lo, hi, d, nsrc, T.cap,
                           (const unsigned long long *)T.key, T.first, T.count, T.depth, T.child, T.bucket, T.counter);
        TCHK(hipGetLastError());
        int nhi = 0;
        TCHK(hipMemcpyAsync(&nhi, T.counter, sizeof(int), hipMemcpyDeviceToHost, st));
        TCHK(hipStreamSynchronize(st));
        if(nhi > T.cap) {
          overflow = true;
          break;
        }
        lo = hi;
        hi = nhi;
        d++;
      }
      if(!overflow) break;
      T.cap *= 2;
      if(attempt == 5) {
        rc = terr(GALIC_ERR_NOMEM, "tree node pool overflow");
        goto done;
      }
    }
  }
  {
    const int nnodes = T.level_lo.back();
    TCHK(B.alloc(&T.mass, nnodes));
    TCHK(B.alloc(&T.sx, nnodes));
    TCHK(B.alloc(&T.sy, nnodes));
    TCHK(B.alloc(&T.sz, nnodes));
    TCHK(B.alloc(&T.len, nnodes));
    if(precision & GALIC_TREE_QUADRUPOLE) {
      TCHK(B.alloc(&T.quad, (size_t)6 * nnodes));
      if(precision & GALIC_TREE_FP32) TCHK(B.alloc(&T.quadf, (size_t)2 * nnodes));
    }
    TCHK(B.alloc(&T.sibling, nnodes));
    TCHK(B.alloc(&T.nextnode, nnodes));
    if((rc = tree_moments_dev(st, T, dlen)) != GALIC_OK) goto done;
  }
  TCHK(hipEventRecord(ee, st));
  TCHK(hipEventSynchronize(ee));
  if(build_ms) TCHK(hipEventElapsedTime(build_ms, eb, ee));
done:
  if(eb) (void)hipEventDestroy(eb);
  if(ee) (void)hipEventDestroy(ee);
  return rc;
}

// The group walk (GALIC_TREE_GROUP, FP32 interactions): the 64 targets of a wave walk the tree
// together.  A node is accepted for the group only if the opening criterion holds at the point
// of the targets' bounding box nearest to its centre of mass, so every target interacts with
// nodes that its own criterion would accept, or with their descendants: more interactions than
// the per-target walk (forcetree_walk.c:12-222), none less accurate.  The wave takes up to 64
// pending entries of an LDS stack at a time (one per lane): accepted nodes, particles and
// opened leaf buckets go to an LDS interaction list, opened nodes push their children.  A full
// list is flushed: every lane adds the list's entries (read at wave-uniform addresses) to its
// own target.  With less stack room than 64 entries could need, one entry is taken at a time
// (a depth-first descent needs at most 7 entries per level).
constexpr int kGStack = 1024, kGList = 1024, kGDeep = 7 * (kMaxDepth + 3);

__device__ __forceinline__ void group_interact(const WalkArgs &A, int e, double px, double py, double pz, float h,
                                               float eps, float &ax, float &ay, float &az, float &pot,
                                               unsigned int &nint) {
  float fac, wp;
  if(e >= 0) {
    float dx, dy, dz, mass;
    int q = -1;
    if(e < A.nsrc) {
      dx = (float)(A.x[e] - px);
      dy = (float)(A.y[e] - py);
      dz = (float)(A.z[e] - pz);
      mass = (float)A.m[e];
    } else {
      q = e - A.nsrc;
      dx = (float)(A.sx[q] - px);
      dy = (float)(A.sy[q] - py);
      dz = (float)(A.sz[q] - pz);
      mass = (float)A.mass[q];
    }
    if(mass == 0) return;
    const float r2 = dx * dx + dy * dy + dz * dz;
    if(q >= 0 && A.quadf) {
      const float4 qa = A.quadf[2 * q], qb = A.quadf[2 * q + 1];
      const float qq[6] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y};
      const float s2 = A.plummer ? r2 + eps * eps : r2;
      if(A.plummer || s2 >= h * h) quad_term(qq, dx, dy, dz, s2, __builtin_amdgcn_rsqf(s2), ax, ay, az, pot);
    }
    if(A.plummer)
      plummer32(mass, r2, eps, fac, wp);
    else
      interact32(mass, r2, h, fac, wp);
    ax += dx * fac;
    ay += dy * fac;
    az += dz * fac;
    pot += wp;
    nint++;
  } else {
    // an opened leaf bucket: its particles one by one
    const int q = ~e, p0 = A.first[q], p1 = p0 + A.count[q];
    for(int p = p0; p < p1; p++) {
      if(A.m[p] == 0) continue;
      const float ex = (float)(A.x[p] - px), ey = (float)(A.y[p] - py), ez = (float)(A.z[p] - pz);
      const float e2 = ex * ex + ey * ey + ez * ez;
      if(A.plummer)
        plummer32((float)A.m[p], e2, eps, fac, wp);
      else
        interact32((float)A.m[p], e2, h, fac, wp);
      ax += ex * fac;
      ay += ey * fac;
      az += ez * fac;
      pot += wp;
      nint++;
    }
  }
}

__device__ __forceinline__ double wave_min(double v) {
  for(int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for(int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__global__ void __launch_bounds__(64) k_walk32g(WalkArgs A) {
  __shared__ int stk[kGStack];
  __shared__ int lst[kGList];
  const int lane = threadIdx.x, g = blockIdx.x;
  long long t;
  if(A.blk_nb > 0) {
    const int nbb = A.blk_nb / 8, s = g / (nbb * nbb), b = g % (nbb * nbb);
    t = ((long long)s * A.blk_nb + (b / nbb) * 8 + lane / 8) * A.blk_nb + (b % nbb) * 8 + lane % 8;
  } else {
    t = (long long)g * 64 + lane;
  }
  const bool live = t < A.ntarget;
  const long long tl = live ? t : (long long)g * 64;  // (dead lanes: the group's first target)
  const double px = A.tx[tl], py = A.ty[tl], pz = A.tz[tl];
  const double bx0 = wave_min(px), by0 = wave_min(py), bz0 = wave_min(pz);
  const double bx1 = wave_max(px), by1 = wave_max(py), bz1 = wave_max(pz);
  const float h = (float)A.h, th2 = (float)(A.theta * A.theta), eps = (float)A.eps;
  float ax = 0, ay = 0, az = 0, pot = 0;
  unsigned int nint = 0;
  if(lane == 0) stk[0] = A.nsrc;  // the root node
  int top = A.nnodes > 0 ? 1 : 0, nl = 0;
  __syncthreads();
  while(top > 0) {
    const int k = (kGStack - top >= 7 * 64 + kGDeep) ? min(64, top) : 1;
    if(nl + k > kGList) {
      for(int i = 0; i < nl; i++)
        group_interact(A, __builtin_amdgcn_readfirstlane(lst[i]), px, py, pz, h, eps, ax, ay, az, pot, nint);
      nl = 0;
    }
    const int e = lane < k ? stk[top - 1 - lane] : -1;
    top -= k;
    __syncthreads();
    int item = 0, open = -1;  // item: the list entry (0: none)
    bool listed = false;
    if(lane < k) {
      if(e < A.nsrc) {
        listed = true;
        item = e;
      } else {
        const int q = e - A.nsrc;
        const double cx = A.sx[q], cy = A.sy[q], cz = A.sz[q];
        const float dx = (float)fmax(fmax(bx0 - cx, cx - bx1), 0.0);
        const float dy = (float)fmax(fmax(by0 - cy, cy - by1), 0.0);
        const float dz = (float)fmax(fmax(bz0 - cz, cz - bz1), 0.0);
        const float l = A.len[q];
        if(l * l > (dx * dx + dy * dy + dz * dz) * th2) {
          if(A.bucket[q]) {
            listed = true;
            item = ~q;
          } else {
            open = q;
          }
        } else {
          listed = true;
          item = e;
        }
      }
    }
    const unsigned long long lm = __ballot(listed);
    if(listed) lst[nl + __builtin_amdgcn_mbcnt_hi((unsigned)(lm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)lm, 0))] = item;
    nl += __popcll(lm);
    for(int c = 0; c < 8; c++) {
      const int ch = open >= 0 ? A.child[(size_t)open * 8 + c] : -1;
      const unsigned long long cm = __ballot(ch >= 0);
      if(ch >= 0) stk[top + __builtin_amdgcn_mbcnt_hi((unsigned)(cm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)cm, 0))] = ch;
      top += __popcll(cm);
    }
    __syncthreads();
  }
  for(int i = 0; i < nl; i++)
    group_interact(A, __builtin_amdgcn_readfirstlane(lst[i]), px, py, pz, h, eps, ax, ay, az, pot, nint);
  if(live) {
    A.ax[t] = (double)ax * A.G;
    A.ay[t] = (double)ay * A.G;
    A.az[t] = (double)az * A.G;
    A.pot[t] = (double)pot * A.G;
    atomicAdd(A.nint, (unsigned long long)nint);
  }
}

// Walk of ntarget targets (device SoA) through T; the results are scaled by G (gravity_tree
// multiplies the walk's sums by G afterwards, gravtree.c:455-463).
// blk_nb > 0: the targets are table nodes (section, z, R) of blk_nb x blk_nb grids, which the group
// walk takes in 8 x 8 blocks.
int tree_walk_dev(const galic_params *p, hipStream_t st, const Tree &T, int ntarget, const double *dtx,
                  const double *dty, const double *dtz, double *dax, double *day, double *daz, double *dpot,
                  int precision, double G, long long *ninteractions, float *walk_ms, int blk_nb = 0) {
  int rc = GALIC_OK;
  DevBufs B;
  unsigned long long *nint = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  WalkArgs W;
  TCHK(B.alloc(&nint, 1));
  TCHK(hipMemsetAsync(nint, 0, sizeof(unsigned long long), st));
  W.nsrc = T.nsrc;
  W.ntarget = ntarget;
  W.tx = dtx;
  W.ty = dty;
  W.tz = dtz;
  W.x = T.x;
  W.y = T.y;
  W.z = T.z;
  W.m = T.m;
  W.nextp = T.nextp;
  W.first = T.first;
  W.count = T.count;
  W.bucket = T.bucket;
  W.sibling = T.sibling;
  W.nextnode = T.nextnode;
  W.mass = T.mass;
  W.sx = T.sx;
  W.sy = T.sy;
  W.sz = T.sz;
  W.len = T.len;
  W.child = T.child;
  W.blk_nb = 0;
  W.quad = (precision & GALIC_TREE_QUADRUPOLE) ? T.quad : nullptr;
  W.quadf = (W.quad && (precision & GALIC_TREE_FP32)) ? T.quadf : nullptr;
  if(W.quad && (precision & GALIC_TREE_FP32) && !W.quadf) {
    rc = terr(GALIC_ERR_STATE, "the tree was built without the FP32 quadrupole copy");
    goto done;
  }
  if((precision & GALIC_TREE_QUADRUPOLE) && !T.quad) {
    rc = terr(GALIC_ERR_STATE, "the tree was built without quadrupole moments");
    goto done;
  }
  W.nnodes = T.level_lo.back();
  W.theta = p->ErrTolTheta;
  W.h = p->ForceSoftening;
  // ForceSoftening is 2.8 times the Softening parameter (init.c:78-84), the Plummer length
  W.plummer = (precision & GALIC_TREE_PLUMMER) ? 1 : 0;
  W.eps = p->ForceSoftening / 2.8;
  W.G = G;
  W.ax = dax;
  W.ay = day;
  W.az = daz;
  W.pot = dpot;
  W.nint = nint;
  TCHK(hipEventCreate(&e0));
  TCHK(hipEventCreate(&e1));
  TCHK(hipEventRecord(e0, st));
  if(ntarget > 0) {
    if(precision & GALIC_TREE_GROUP) {
      const bool blk = blk_nb > 0 && blk_nb % 8 == 0 && (long long)ntarget % ((long long)blk_nb * blk_nb) == 0;
      W.blk_nb = blk ? blk_nb : 0;
      hipLaunchKernelGGL(k_walk32g, dim3((unsigned int)(((long long)ntarget + 63) / 64)), dim3(64), 0, st, W);
    } else if(!(precision & GALIC_TREE_FP32)) {
      hipLaunchKernelGGL(k_walk64, dim3(nb256(ntarget)), dim3(256), 0, st, W);
    } else {
      hipLaunchKernelGGL(k_walk32, dim3(nb256(ntarget)), dim3(256), 0, st, W);
    }
    TCHK(hipGetLastError());
  }
  TCHK(hipEventRecord(e1, st));
  TCHK(hipEventSynchronize(e1));
  if(walk_ms) TCHK(hipEventElapsedTime(walk_ms, e0, e1));
  {
    unsigned long long ni = 0;
    TCHK(hipMemcpy(&ni, nint, sizeof(ni), hipMemcpyDeviceToHost));
    if(ninteractions) *ninteractions = (long long)ni;
  }
done:
  if(e0) (void)hipEventDestroy(e0);
  if(e1) (void)hipEventDestroy(e1);
  return rc;
}

// Build the tree of nsrc sources (device SoA inputs) and walk ntarget targets.
int tree_forces_dev(const galic_params *p, hipStream_t st, int nsrc, const double *dpx, const double *dpy,
                    const double *dpz, const double *dpm, int ntarget, const double *dtx, const double *dty,
                    const double *dtz, double corner[3], double dlen, double *dax, double *day, double *daz,
                    double *dpot, int precision, long long *ninteractions, float *walk_ms, int blk_nb = 0) {
  DevBufs B;
  Tree T;
  float build_ms = 0, wms = 0;
  long long ni = 0;
  int rc = tree_build_dev(p, st, nsrc, dpx, dpy, dpz, dpm, corner, dlen, precision, B, T, nullptr, &build_ms);
  if(rc != GALIC_OK) return rc;
  rc = tree_walk_dev(p, st, T, ntarget, dtx, dty, dtz, dax, day, daz, dpot, precision, p->G, &ni, &wms, blk_nb);
  if(rc != GALIC_OK) return rc;
  if(ninteractions) *ninteractions = ni;
  if(walk_ms) *walk_ms = wms;
  g_last_build_ms = build_ms;
  g_last_walk_ms = wms;
  g_last_interactions = (double)ni;
  return GALIC_OK;
}

// domain_box.c:20-65 over sources and targets
void domain_extent(int nsrc, const double *spos, int ntarget, const double *tpos, double corner[3], double *len) {
  double xmin[3] = {1e30, 1e30, 1e30}, xmax[3] = {-1e30, -1e30, -1e30}, l = 0;
  for(int i = 0; i < nsrc + ntarget; i++) {
    const double *x = i < nsrc ? &spos[3 * (size_t)i] : &tpos[3 * (size_t)(i - nsrc)];
    for(int j = 0; j < 3; j++) {
      if(xmin[j] > x[j]) xmin[j] = x[j];
      if(xmax[j] < x[j]) xmax[j] = x[j];
    }
  }
  for(int j = 0; j < 3; j++)
    if(xmax[j] - xmin[j] > l) l = xmax[j] - xmin[j];
  l *= 1.00001;
  for(int j = 0; j < 3; j++) corner[j] = 0.5 * (xmin[j] + xmax[j]) - 0.5 * l;
  *len = l;
}

int upload_soa(hipStream_t st, int n, const double *aos, double *dx, double *dy, double *dz) {
  std::vector<double> t(std::max(n, 1));
  double *d[3] = {dx, dy, dz};
  for(int k = 0; k < 3; k++) {
    for(int i = 0; i < n; i++) t[i] = aos[3 * (size_t)i + k];
    if(n > 0 && hipMemcpyAsync(d[k], t.data(), sizeof(double) * n, hipMemcpyHostToDevice, st) != hipSuccess)
      return GALIC_ERR_HIP;
    if(hipStreamSynchronize(st) != hipSuccess) return GALIC_ERR_HIP;
  }
  return GALIC_OK;
}

// Node radii R_j = Rmin (Fac^j - 1) of the force grid (grid.c:680-704) and the section
// angles' cosines and sines, as galic_tree_fill computes them.
void fill_nodes(const galic_params *p, std::vector<double> &rj, std::vector<double> &cs, std::vector<double> &sn) {
  const int nb = p->FG_Nbin;
  rj.resize(nb);
  cs.resize(GALIC_FG_SECTIONS);
  sn.resize(GALIC_FG_SECTIONS);
  for(int j = 0; j < nb; j++) rj[j] = p->FG_Rmin * (pow(p->FG_Fac, j) - 1.0);
  for(int s = 0; s < GALIC_FG_SECTIONS; s++) {
    double phi = 2 * M_PI / GALIC_FG_SECTIONS * s;
    cs[s] = cos(phi);
    sn[s] = sin(phi);
  }
}

// ------------------------------------------------ forcetree drop-ins' state --

// P[] of the forcetree drop-ins (allvars.h:844-868), caller-owned as the reference's P[]:
// positions, masses, types, and the GravAccel and Potential the walks write.
struct BoundParticles {
  int n = 0;
  double *pos = nullptr, *mass = nullptr, *acc = nullptr, *pot = nullptr;
  int *type = nullptr;
} g_P;

// The tree of the last force_treebuild, kept on the device for force_treeevaluate and
// force_update_node_recursive.
struct BoundTree {
  bool built = false;
  int device = 0;
  int npart = 0;                 // the build's npart: the root node's index (the reference's Tree_MaxPart)
  hipStream_t st = nullptr;
  DevBufs *B = nullptr;
  Tree T;
  int *sorted = nullptr;         // device: sorted source -> source index
  std::vector<int> src_of;       // source index -> index in P[]
  std::vector<int> targets;      // TargetList: the Type-5 particles (gravtree.c:115-119)
  double corner[3] = {0, 0, 0}, len = 0;
  galic_params p;
  double *tbuf = nullptr;        // device: one target's position, acceleration and potential
} g_T;

void tree_release() {
  if(g_T.B) {
    (void)hipSetDevice(g_T.device);
    delete g_T.B;
  }
  if(g_T.st) (void)hipStreamDestroy(g_T.st);
  g_T = BoundTree();
}

void tree_dropin_fail(const char *what) {
  // the reference terminates on errors in the tree code (allvars.h:129)
  fprintf(stderr, "galic forcetree drop-in %s failed: %s\n", what, galic_last_error());
  exit(1);
}

// Builds the bound tree over P[0..npart) without the Type-5 particles (forcetree.c:331);
// the domain cube covers every one of them (domain_box.c:20-65).  Returns the node count.
int tree_build_bound(int npart, int precision) {
  galic_ctx *ctx = galic::bound_ctx();
  if(!ctx) return terr(GALIC_ERR_STATE, "no context bound (galic_bind): the tree takes its parameters from it");
  if(!g_P.pos || npart < 0 || npart > g_P.n)
    return terr(GALIC_ERR_ARG, "force_treebuild: npart %d beyond the %d bound particles (galic_bind_tree_particles)",
                npart, g_P.n);
  tree_release();
  const galic::Ctx &c = galic::ctx_state(ctx);
  g_T.p = c.p;
  g_T.device = c.device;
  g_T.npart = npart;
  int rc = GALIC_OK;
  std::vector<double> sx, sy, sz, sm, tp;
  std::vector<double> spos;
  for(int i = 0; i < npart; i++) {
    if(g_P.type[i] == 5) {
      g_T.targets.push_back(i);
      tp.insert(tp.end(), g_P.pos + 3 * (size_t)i, g_P.pos + 3 * (size_t)i + 3);
    } else {
      g_T.src_of.push_back(i);
      spos.insert(spos.end(), g_P.pos + 3 * (size_t)i, g_P.pos + 3 * (size_t)i + 3);
      sm.push_back(g_P.mass[i]);
    }
  }
  const int nsrc = (int)g_T.src_of.size();
  double *px, *py, *pz, *pm;
  domain_extent(nsrc, spos.data(), (int)g_T.targets.size(), tp.data(), g_T.corner, &g_T.len);
  g_T.B = new DevBufs();
  TCHK(hipSetDevice(g_T.device));
  TCHK(hipStreamCreate(&g_T.st));
  TCHK(g_T.B->alloc(&px, nsrc));
  TCHK(g_T.B->alloc(&py, nsrc));
  TCHK(g_T.B->alloc(&pz, nsrc));
  TCHK(g_T.B->alloc(&pm, nsrc));
  TCHK(g_T.B->alloc(&g_T.tbuf, 8));
  if((rc = upload_soa(g_T.st, nsrc, spos.data(), px, py, pz)) != GALIC_OK) goto done;
  if(nsrc > 0) TCHK(hipMemcpy(pm, sm.data(), sizeof(double) * nsrc, hipMemcpyHostToDevice));
  rc = tree_build_dev(&g_T.p, g_T.st, nsrc, px, py, pz, pm, g_T.corner, g_T.len, precision, *g_T.B, g_T.T,
                      &g_T.sorted, nullptr);
  if(rc == GALIC_OK) g_T.built = true;
done:
  if(rc != GALIC_OK) tree_release();
  return rc == GALIC_OK ? g_T.T.level_lo.back() : rc;
}

// Walks n targets (positions pos, 3 each) of the bound tree; G scales the results.
int tree_walk_bound(int n, const double *pos, double *acc, double *pot, double G, long long *nint) {
  int rc = GALIC_OK;
  DevBufs B;
  double *tx, *ty, *tz, *ax, *ay, *az, *pt;
  std::vector<double> h;
  TCHK(hipSetDevice(g_T.device));
  TCHK(B.alloc(&tx, n)); TCHK(B.alloc(&ty, n)); TCHK(B.alloc(&tz, n));
  TCHK(B.alloc(&ax, n)); TCHK(B.alloc(&ay, n)); TCHK(B.alloc(&az, n)); TCHK(B.alloc(&pt, n));
  if((rc = upload_soa(g_T.st, n, pos, tx, ty, tz)) != GALIC_OK) goto done;
  if((rc = tree_walk_dev(&g_T.p, g_T.st, g_T.T, n, tx, ty, tz, ax, ay, az, pt, 0, G, nint, nullptr)) != GALIC_OK)
    goto done;
  h.resize(std::max(n, 1));
  {
    double *src[3] = {ax, ay, az};
    for(int k = 0; k < 3; k++) {
      if(n > 0) TCHK(hipMemcpy(h.data(), src[k], sizeof(double) * n, hipMemcpyDeviceToHost));
      for(int i = 0; i < n; i++) acc[3 * (size_t)i + k] = h[i];
    }
    if(n > 0) TCHK(hipMemcpy(pot, pt, sizeof(double) * n, hipMemcpyDeviceToHost));
  }
done:
  return rc;
}

}  // namespace

extern "C" {

int galic_bind_tree_particles(int n, double *pos, double *mass, int *type, double *grav_accel, double *potential) {
  if(n < 0 || (n > 0 && (!pos || !mass || !type || !grav_accel || !potential)))
    return terr(GALIC_ERR_ARG, "galic_bind_tree_particles: bad arguments");
  tree_release();
  g_P.n = n;
  g_P.pos = pos;
  g_P.mass = mass;
  g_P.type = type;
  g_P.acc = grav_accel;
  g_P.pot = potential;
  return GALIC_OK;
}

int force_treebuild(int npart, int optimized_domain_mapping) {
  (void)optimized_domain_mapping;  // one GPU: no domain mapping to optimise
  const int r = tree_build_bound(npart, 0);
  if(r < 0) tree_dropin_fail("force_treebuild");
  return r;
}

void force_update_node_recursive(int no, int sib, int father, int *last) {
  if(!g_T.built) tree_dropin_fail("force_update_node_recursive (no tree built)");
  if(no != g_T.npart || sib != -1 || father != -1) {
    terr(GALIC_ERR_ARG, "node %d: the device tree numbers its own nodes; only the root (%d, sibling -1, father -1, "
         "as forcetree.c:86 calls it) is addressed", no, g_T.npart);
    tree_dropin_fail("force_update_node_recursive");
  }
  (void)hipSetDevice(g_T.device);
  if(tree_moments_dev(g_T.st, g_T.T, g_T.len) != GALIC_OK || hipStreamSynchronize(g_T.st) != hipSuccess)
    tree_dropin_fail("force_update_node_recursive");
  if(last) {
    // the walk order ends at the last source in key order (children are threaded in order)
    int s = -1;
    const int nsrc = g_T.T.nsrc;
    if(nsrc > 0 && hipMemcpy(&s, g_T.sorted + (nsrc - 1), sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      tree_dropin_fail("force_update_node_recursive");
    *last = (s >= 0) ? g_T.src_of[s] : -1;
  }
}

int force_treeevaluate(int target, int mode, int thread_id) {
  (void)thread_id;
  if(!g_T.built) tree_dropin_fail("force_treeevaluate (no tree built)");
  if(mode != 0) {
    terr(GALIC_ERR_ARG, "mode %d: imported targets of other tasks have no counterpart on one GPU", mode);
    tree_dropin_fail("force_treeevaluate");
  }
  if(target < 0 || target >= (int)g_T.targets.size()) {
    terr(GALIC_ERR_ARG, "target %d beyond the %zu Type-5 particles", target, g_T.targets.size());
    tree_dropin_fail("force_treeevaluate");
  }
  // TargetList[target] (forcetree_walk.c:28-41); the sums are left without G, as the walk
  // leaves them (gravity_tree multiplies afterwards)
  const int i = g_T.targets[target];
  long long nint = 0;
  if(tree_walk_bound(1, g_P.pos + 3 * (size_t)i, g_P.acc + 3 * (size_t)i, g_P.pot + i, 1.0, &nint) != GALIC_OK)
    tree_dropin_fail("force_treeevaluate");
  return (int)nint;
}

void gravity(void) {
  // gravtree.c:35-50 and 88-467: build, walk every Type-5 particle, multiply by G, free
  if(tree_build_bound(g_P.n, 0) < 0) tree_dropin_fail("gravity");
  const int nt = (int)g_T.targets.size();
  std::vector<double> tp(3 * (size_t)nt), acc(3 * (size_t)nt), pot(nt);
  for(int k = 0; k < nt; k++)
    for(int j = 0; j < 3; j++) tp[3 * (size_t)k + j] = g_P.pos[3 * (size_t)g_T.targets[k] + j];
  if(tree_walk_bound(nt, tp.data(), acc.data(), pot.data(), g_T.p.G, nullptr) != GALIC_OK) tree_dropin_fail("gravity");
  for(int k = 0; k < nt; k++) {
    const int i = g_T.targets[k];
    for(int j = 0; j < 3; j++) g_P.acc[3 * (size_t)i + j] = acc[3 * (size_t)k + j];
    g_P.pot[i] = pot[k];
  }
  tree_release();
}

int galic_set_tree_sources(galic_ctx *ctx, int nsrc, const double *spos, const double *smass) {
  if(!ctx || nsrc < 1 || !spos || !smass) return terr(GALIC_ERR_ARG, "galic_set_tree_sources: bad arguments");
  galic::Ctx &c = galic::ctx_state(ctx);
  int rc = GALIC_OK;
  TCHK(hipSetDevice(c.device));
  if(c.tsrc) (void)hipFree(c.tsrc);
  c.tsrc = nullptr;
  c.tsrc_n = 0;
  TCHK(hipMalloc(&c.tsrc, sizeof(double) * 4 * (size_t)nsrc));
  c.tsrc_n = nsrc;
  for(int j = 0; j < 3; j++) {
    c.tsrc_lo[j] = 1e30;
    c.tsrc_hi[j] = -1e30;
  }
  for(int i = 0; i < nsrc; i++)
    for(int j = 0; j < 3; j++) {
      const double x = spos[3 * (size_t)i + j];
      if(c.tsrc_lo[j] > x) c.tsrc_lo[j] = x;
      if(c.tsrc_hi[j] < x) c.tsrc_hi[j] = x;
    }
  if((rc = upload_soa(c.stream, nsrc, spos, c.tsrc, c.tsrc + nsrc, c.tsrc + 2 * (size_t)nsrc)) != GALIC_OK) goto done;
  TCHK(hipMemcpy(c.tsrc + 3 * (size_t)nsrc, smass, sizeof(double) * nsrc, hipMemcpyHostToDevice));
done:
  return rc;
}

int galic_retabulate(galic_ctx *ctx, int precision, long long *ninteractions) {
  if(!ctx) return terr(GALIC_ERR_ARG, "null context");
  if((precision & ~15) || ((precision & GALIC_TREE_GROUP) && !(precision & GALIC_TREE_FP32)))
    return terr(GALIC_ERR_ARG, "mode: GALIC_TREE_FP32 | GALIC_TREE_QUADRUPOLE | GALIC_TREE_PLUMMER | "
                               "GALIC_TREE_GROUP (with GALIC_TREE_FP32)");
  galic::Ctx &c = galic::ctx_state(ctx);
  if(!c.tsrc || c.tsrc_n < 1) return terr(GALIC_ERR_STATE, "no tree sources (galic_set_tree_sources)");
  const galic_params *p = &c.p;
  const int nb = p->FG_Nbin, nsrc = c.tsrc_n;
  const size_t nn = (size_t)nb * nb, ns = GALIC_FG_SECTIONS * nn;
  std::vector<double> rj, cs, sn;
  fill_nodes(p, rj, cs, sn);
  // the domain over sources and targets, as domain_extent() forms it for galic_tree_fill
  double xmin[3], xmax[3], corner[3], len = 0;
  for(int j = 0; j < 3; j++) {
    xmin[j] = c.tsrc_lo[j];
    xmax[j] = c.tsrc_hi[j];
  }
  for(int s = 0; s < GALIC_FG_SECTIONS; s++)
    for(int j = 0; j < nb; j++) {
      const double x = rj[j] * cs[s], y = rj[j] * sn[s];
      xmin[0] = std::min(xmin[0], x);
      xmax[0] = std::max(xmax[0], x);
      xmin[1] = std::min(xmin[1], y);
      xmax[1] = std::max(xmax[1], y);
    }
  for(int k = 0; k < nb; k++) {
    xmin[2] = std::min(xmin[2], rj[k]);
    xmax[2] = std::max(xmax[2], rj[k]);
  }
  for(int j = 0; j < 3; j++)
    if(xmax[j] - xmin[j] > len) len = xmax[j] - xmin[j];
  len *= 1.00001;
  for(int j = 0; j < 3; j++) corner[j] = 0.5 * (xmin[j] + xmax[j]) - 0.5 * len;
  int rc = GALIC_OK;
  DevBufs B;
  Tree T;
  double *drj, *dcs, *dsn, *tx, *ty, *tz, *ax, *ay, *az, *pt;
  long long ni = 0;
  float bms = 0, wms = 0;
  TCHK(hipSetDevice(c.device));
  TCHK(B.alloc(&drj, nb)); TCHK(B.alloc(&dcs, GALIC_FG_SECTIONS)); TCHK(B.alloc(&dsn, GALIC_FG_SECTIONS));
  TCHK(B.alloc(&tx, ns)); TCHK(B.alloc(&ty, ns)); TCHK(B.alloc(&tz, ns));
  TCHK(B.alloc(&ax, ns)); TCHK(B.alloc(&ay, ns)); TCHK(B.alloc(&az, ns)); TCHK(B.alloc(&pt, ns));
  TCHK(hipMemcpyAsync(drj, rj.data(), sizeof(double) * nb, hipMemcpyHostToDevice, c.stream));
  TCHK(hipMemcpyAsync(dcs, cs.data(), sizeof(double) * GALIC_FG_SECTIONS, hipMemcpyHostToDevice, c.stream));
  TCHK(hipMemcpyAsync(dsn, sn.data(), sizeof(double) * GALIC_FG_SECTIONS, hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_fill_targets, dim3(nb256(ns)), dim3(256), 0, c.stream, nb, (const double *)drj,
                     (const double *)dcs, (const double *)dsn, tx, ty, tz);
  TCHK(hipGetLastError());
  if((rc = tree_build_dev(p, c.stream, nsrc, c.tsrc, c.tsrc + nsrc, c.tsrc + 2 * (size_t)nsrc,
                          c.tsrc + 3 * (size_t)nsrc, corner, len, precision, B, T, nullptr, &bms)) != GALIC_OK)
    goto done;
  if((rc = tree_walk_dev(p, c.stream, T, (int)ns, tx, ty, tz, ax, ay, az, pt, precision, p->G, &ni, &wms, nb)) !=
     GALIC_OK)
    goto done;
  // installed as the context's tables: FG_DPotDR, FG_DPotDz, FG_Pot and the FP32 copy
  hipLaunchKernelGGL(k_project, dim3(nb256(nn)), dim3(256), 0, c.stream, nb, (const double *)ax, (const double *)ay,
                     (const double *)az, (const double *)pt, (const double *)dcs, (const double *)dsn, c.FG_DPotDR,
                     c.FG_DPotDz, c.FG_Pot);
  hipLaunchKernelGGL(k_acc_f, dim3(nb256(nn)), dim3(256), 0, c.stream, (int)nn, (const double *)c.FG_DPotDR,
                     (const double *)c.FG_DPotDz, c.FG_Acc_f);
  TCHK(hipGetLastError());
  TCHK(hipStreamSynchronize(c.stream));
  if(ninteractions) *ninteractions = ni;
  g_last_build_ms = bms;
  g_last_walk_ms = wms;
  g_last_interactions = (double)ni;
done:
  return rc;
}

int galic_tree_forces(const galic_params *p, int device, int nsrc, const double *spos, const double *smass,
                      int ntarget, const double *tpos, double *acc, double *pot, int precision,
                      long long *ninteractions) {
  if(!p || nsrc < 0 || ntarget < 0 || (nsrc > 0 && (!spos || !smass)) || (ntarget > 0 && (!tpos || !acc || !pot)))
    return terr(GALIC_ERR_ARG, "bad tree arguments");
  if((precision & ~15) || ((precision & GALIC_TREE_GROUP) && !(precision & GALIC_TREE_FP32)))
    return terr(GALIC_ERR_ARG, "mode: GALIC_TREE_FP32 | GALIC_TREE_QUADRUPOLE | GALIC_TREE_PLUMMER | "
                               "GALIC_TREE_GROUP (with GALIC_TREE_FP32)");
  int rc = GALIC_OK;
  hipStream_t st = nullptr;
  DevBufs B;
  double *px, *py, *pz, *pm, *tx, *ty, *tz, *ax, *ay, *az, *pt;
  double corner[3], len;
  float ms = 0;
  std::vector<double> h;
  TCHK(hipSetDevice(device));
  TCHK(hipStreamCreate(&st));
  domain_extent(nsrc, spos, ntarget, tpos, corner, &len);
  TCHK(B.alloc(&px, nsrc)); TCHK(B.alloc(&py, nsrc)); TCHK(B.alloc(&pz, nsrc)); TCHK(B.alloc(&pm, nsrc));
  TCHK(B.alloc(&tx, ntarget)); TCHK(B.alloc(&ty, ntarget)); TCHK(B.alloc(&tz, ntarget));
  TCHK(B.alloc(&ax, ntarget)); TCHK(B.alloc(&ay, ntarget)); TCHK(B.alloc(&az, ntarget)); TCHK(B.alloc(&pt, ntarget));
  if((rc = upload_soa(st, nsrc, spos, px, py, pz)) != GALIC_OK) goto done;
  if((rc = upload_soa(st, ntarget, tpos, tx, ty, tz)) != GALIC_OK) goto done;
  if(nsrc > 0) TCHK(hipMemcpy(pm, smass, sizeof(double) * nsrc, hipMemcpyHostToDevice));
  rc = tree_forces_dev(p, st, nsrc, px, py, pz, pm, ntarget, tx, ty, tz, corner, len, ax, ay, az, pt, precision,
                       ninteractions, &ms);
  if(rc != GALIC_OK) goto done;
  h.resize(std::max(ntarget, 1));
  {
    double *src[3] = {ax, ay, az};
    for(int k = 0; k < 3; k++) {
      if(ntarget > 0) TCHK(hipMemcpy(h.data(), src[k], sizeof(double) * ntarget, hipMemcpyDeviceToHost));
      for(int i = 0; i < ntarget; i++) acc[3 * (size_t)i + k] = h[i];
    }
    if(ntarget > 0) TCHK(hipMemcpy(pot, pt, sizeof(double) * ntarget, hipMemcpyDeviceToHost));
  }
done:
  if(st) (void)hipStreamDestroy(st);
  return rc;
}

int galic_tree_fill(const galic_params *p, int device, int nsrc, const double *spos, const double *smass,
                    double *pot, double *dpotdr, double *dpotdz, int precision, long long *ninteractions) {
  if(!p || nsrc < 0 || (nsrc > 0 && (!spos || !smass)) || !pot || !dpotdr || !dpotdz || p->FG_Nbin < 2)
    return terr(GALIC_ERR_ARG, "bad table-fill arguments");
  if((precision & ~15) || ((precision & GALIC_TREE_GROUP) && !(precision & GALIC_TREE_FP32)))
    return terr(GALIC_ERR_ARG, "mode: GALIC_TREE_FP32 | GALIC_TREE_QUADRUPOLE | GALIC_TREE_PLUMMER | "
                               "GALIC_TREE_GROUP (with GALIC_TREE_FP32)");
  const int nb = p->FG_Nbin;
  const size_t ns = (size_t)GALIC_FG_SECTIONS * nb * nb;
  if(ns > 0x7fffffff) return terr(GALIC_ERR_ARG, "FG_Nbin too large");
  // targets of grid.c:680-704: (R_j cos phi_s, R_j sin phi_s, z_k)
  std::vector<double> tpos(3 * ns), cs(GALIC_FG_SECTIONS), sn(GALIC_FG_SECTIONS);
  for(int s = 0; s < GALIC_FG_SECTIONS; s++) {
    double phi = 2 * M_PI / GALIC_FG_SECTIONS * s;
    cs[s] = cos(phi);
    sn[s] = sin(phi);
  }
  size_t n = 0;
  for(int s = 0; s < GALIC_FG_SECTIONS; s++)
    for(int k = 0; k < nb; k++)
      for(int j = 0; j < nb; j++, n++) {
        double r1 = p->FG_Rmin * (pow(p->FG_Fac, j) - 1.0);
        double z1 = p->FG_Rmin * (pow(p->FG_Fac, k) - 1.0);
        tpos[3 * n] = r1 * cs[s];
        tpos[3 * n + 1] = r1 * sn[s];
        tpos[3 * n + 2] = z1;
      }
  int rc = GALIC_OK;
  hipStream_t st = nullptr;
  DevBufs B;
  double *px, *py, *pz, *pm, *tx, *ty, *tz, *ax, *ay, *az, *pt, *dcs, *dsn, *r0, *r1, *r2;
  double corner[3], len;
  float ms = 0;
  TCHK(hipSetDevice(device));
  TCHK(hipStreamCreate(&st));
  domain_extent(nsrc, spos, (int)ns, tpos.data(), corner, &len);
  TCHK(B.alloc(&px, nsrc)); TCHK(B.alloc(&py, nsrc)); TCHK(B.alloc(&pz, nsrc)); TCHK(B.alloc(&pm, nsrc));
  TCHK(B.alloc(&tx, ns)); TCHK(B.alloc(&ty, ns)); TCHK(B.alloc(&tz, ns));
  TCHK(B.alloc(&ax, ns)); TCHK(B.alloc(&ay, ns)); TCHK(B.alloc(&az, ns)); TCHK(B.alloc(&pt, ns));
  TCHK(B.alloc(&dcs, GALIC_FG_SECTIONS)); TCHK(B.alloc(&dsn, GALIC_FG_SECTIONS));
  TCHK(B.alloc(&r0, (size_t)nb * nb)); TCHK(B.alloc(&r1, (size_t)nb * nb)); TCHK(B.alloc(&r2, (size_t)nb * nb));
  if((rc = upload_soa(st, nsrc, spos, px, py, pz)) != GALIC_OK) goto done;
  if((rc = upload_soa(st, (int)ns, tpos.data(), tx, ty, tz)) != GALIC_OK) goto done;
  if(nsrc > 0) TCHK(hipMemcpy(pm, smass, sizeof(double) * nsrc, hipMemcpyHostToDevice));
  TCHK(hipMemcpy(dcs, cs.data(), sizeof(double) * GALIC_FG_SECTIONS, hipMemcpyHostToDevice));
  TCHK(hipMemcpy(dsn, sn.data(), sizeof(double) * GALIC_FG_SECTIONS, hipMemcpyHostToDevice));
  rc = tree_forces_dev(p, st, nsrc, px, py, pz, pm, (int)ns, tx, ty, tz, corner, len, ax, ay, az, pt, precision,
                       ninteractions, &ms, nb);
  if(rc != GALIC_OK) goto done;
  hipLaunchKernelGGL(k_project, dim3(nb256((size_t)nb * nb)), dim3(256), 0, st, nb, (const double *)ax,
                     (const double *)ay, (const double *)az, (const double *)pt, (const double *)dcs,
                     (const double *)dsn, r1, r2, r0);
  TCHK(hipGetLastError());
  TCHK(hipStreamSynchronize(st));
  TCHK(hipMemcpy(dpotdr, r1, sizeof(double) * nb * nb, hipMemcpyDeviceToHost));
  TCHK(hipMemcpy(dpotdz, r2, sizeof(double) * nb * nb, hipMemcpyDeviceToHost));
  TCHK(hipMemcpy(pot, r0, sizeof(double) * nb * nb, hipMemcpyDeviceToHost));
done:
  if(st) (void)hipStreamDestroy(st);
  return rc;
}

int galic_direct_forces(const galic_params *p, int device, int nsrc, const double *spos, const double *smass,
                        int ntarget, const double *tpos, double *acc, double *pot) {
  return galic_direct_forces_mode(p, device, nsrc, spos, smass, ntarget, tpos, acc, pot, 0);
}

int galic_direct_forces_mode(const galic_params *p, int device, int nsrc, const double *spos, const double *smass,
                             int ntarget, const double *tpos, double *acc, double *pot, int mode) {
  if(!p || nsrc < 0 || ntarget < 0 || (nsrc > 0 && (!spos || !smass)) || (ntarget > 0 && (!tpos || !acc || !pot)))
    return terr(GALIC_ERR_ARG, "bad direct-summation arguments");
  if(mode != 0 && mode != GALIC_TREE_PLUMMER) return terr(GALIC_ERR_ARG, "mode: 0 or GALIC_TREE_PLUMMER");
  int rc = GALIC_OK;
  hipStream_t st = nullptr;
  DevBufs B;
  double *px, *py, *pz, *pm, *tx, *ty, *tz, *ax, *ay, *az, *pt;
  std::vector<double> h;
  TCHK(hipSetDevice(device));
  TCHK(hipStreamCreate(&st));
  TCHK(B.alloc(&px, nsrc)); TCHK(B.alloc(&py, nsrc)); TCHK(B.alloc(&pz, nsrc)); TCHK(B.alloc(&pm, nsrc));
  TCHK(B.alloc(&tx, ntarget)); TCHK(B.alloc(&ty, ntarget)); TCHK(B.alloc(&tz, ntarget));
  TCHK(B.alloc(&ax, ntarget)); TCHK(B.alloc(&ay, ntarget)); TCHK(B.alloc(&az, ntarget)); TCHK(B.alloc(&pt, ntarget));
  if((rc = upload_soa(st, nsrc, spos, px, py, pz)) != GALIC_OK) goto done;
  if((rc = upload_soa(st, ntarget, tpos, tx, ty, tz)) != GALIC_OK) goto done;
  if(nsrc > 0) TCHK(hipMemcpy(pm, smass, sizeof(double) * nsrc, hipMemcpyHostToDevice));
  if(ntarget > 0) {
    hipLaunchKernelGGL(galic::tree::k_direct64, dim3(nb256(ntarget)), dim3(256), 0, st, nsrc, (const double *)px,
                       (const double *)py, (const double *)pz, (const double *)pm, ntarget, (const double *)tx,
                       (const double *)ty, (const double *)tz, p->ForceSoftening, mode == GALIC_TREE_PLUMMER ? 1 : 0,
                       p->ForceSoftening / 2.8, p->G, ax, ay, az, pt);
    TCHK(hipGetLastError());
    TCHK(hipStreamSynchronize(st));
    h.resize(ntarget);
    double *src[3] = {ax, ay, az};
    for(int k = 0; k < 3; k++) {
      TCHK(hipMemcpy(h.data(), src[k], sizeof(double) * ntarget, hipMemcpyDeviceToHost));
      for(int i = 0; i < ntarget; i++) acc[3 * (size_t)i + k] = h[i];
    }
    TCHK(hipMemcpy(pot, pt, sizeof(double) * ntarget, hipMemcpyDeviceToHost));
  }
done:
  if(st) (void)hipStreamDestroy(st);
  return rc;
}

int galic_tree_last_timings(double *out3) {
  if(!out3) return terr(GALIC_ERR_ARG, "null argument");
  out3[0] = g_last_build_ms;
  out3[1] = g_last_walk_ms;
  out3[2] = g_last_interactions;
  return GALIC_OK;
}

}  // extern "C"
