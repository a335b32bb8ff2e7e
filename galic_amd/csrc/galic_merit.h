// galic_merit.h -- the change of one orbit-averaged merit term on the fixed
// density cut, shared by the FP64 and FP32 merit kernels.
//
// For field f the term is sum over cut nodes c of |R_c + D_c - T_c| * w_c
// (calc_stack_difference / _mod, grid.c:1086-1151), D the trial's change of the
// response.  Its change is sum_c w_c (|b_c + D_c| - |b_c|), b = R - T, over the
// nodes with D_c != 0: the fine cells of the orbits' bounding box and their
// ancestors.  Node values of D are built level by level in LDS in the child
// order of smooth_stack (grid.c:1256-1260), only over the box's ancestors.
#pragma once
#include <hip/hip_runtime.h>

#include "galic_ctx.h"

namespace galic {

struct CutTerm {
  const double *R, *T;   // response and target stacks of the field
  const double *den;     // min(dist, R200/20) on cut nodes, 0 elsewhere
  const double *ratio;   // Tm/E_x on cut nodes of an S term (0: no term), null for the mass term
};

__device__ __forceinline__ double cut_phi(const CutTerm &C, int node, double d) {
  if(d == 0) return 0.0;
  double den = C.den[node];
  if(den == 0) return 0.0;
  double wgt = C.ratio ? C.ratio[node] : 1.0;
  if(wgt == 0) return 0.0;
  double b = C.R[node] - C.T[node];
  return (fabs(b + d) - fabs(b)) * wgt / den;
}

// Per-thread partial of the change; all threads of the block must call it.
// loadD(cell) returns D at fine cell `cell`; cells outside the box read as 0.
// cs: LDS of at least GALIC_NSTACK(L) - 4^L doubles.
template <typename LoadD>
__device__ double cut_sum_field(const CutTerm &C, LoadD loadD, int r0, int r1, int z0, int z1, int L, double *cs) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nb = 1 << L;
  const int offL = GALIC_STACKOFFSET(L, 0, 0);
  double phi = 0;
  if(r1 < r0 || z1 < z0) {
    __syncthreads();
    return 0.0;
  }
  // level L-1 from the fine level; fine cut nodes are evaluated on the way
  int a0 = r0 >> 1, a1 = r1 >> 1, c0 = z0 >> 1, c1 = z1 >> 1;
  int na = a1 - a0 + 1, nc = c1 - c0 + 1;
  for(int k = tid; k < na * nc; k += nt) {
    int pz = c0 + k / na, pr = a0 + k % na;
    int cz[4] = {2 * pz, 2 * pz + 1, 2 * pz, 2 * pz + 1};
    int cr[4] = {2 * pr, 2 * pr, 2 * pr + 1, 2 * pr + 1};
    double d[4];
    for(int c = 0; c < 4; c++) {
      int cell = cz[c] * nb + cr[c];
      d[c] = loadD(cell);
      phi += cut_phi(C, offL + cell, d[c]);
    }
    double s = d[0] + d[1] + d[2] + d[3];
    int node = GALIC_STACKOFFSET(L - 1, pz, pr);
    cs[node] = s;
    phi += cut_phi(C, node, s);
  }
  __syncthreads();
  // coarser levels over the box's ancestors; children outside the previous
  // region hold stale LDS and are read as 0
  for(int l = L - 2; l >= 0; l--) {
    int b0 = a0 >> 1, b1 = a1 >> 1, e0 = c0 >> 1, e1 = c1 >> 1;
    int nbx = b1 - b0 + 1, nez = e1 - e0 + 1;
    for(int k = tid; k < nbx * nez; k += nt) {
      int i = e0 + k / nbx, j = b0 + k % nbx;
      double s = 0;
      const int ci[4] = {2 * i, 2 * i + 1, 2 * i, 2 * i + 1};
      const int cj[4] = {2 * j, 2 * j, 2 * j + 1, 2 * j + 1};
      for(int c = 0; c < 4; c++) {
        bool in = ci[c] >= c0 && ci[c] <= c1 && cj[c] >= a0 && cj[c] <= a1;
        s += in ? cs[GALIC_STACKOFFSET(l + 1, ci[c], cj[c])] : 0.0;
      }
      int node = GALIC_STACKOFFSET(l, i, j);
      cs[node] = s;
      phi += cut_phi(C, node, s);
    }
    a0 = b0;
    a1 = b1;
    c0 = e0;
    c1 = e1;
    __syncthreads();
  }
  return phi;
}

}  // namespace galic
