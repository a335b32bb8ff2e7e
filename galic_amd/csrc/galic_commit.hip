// galic_commit.hip -- adding response logs to the fixed-point Loc grids of the FP32 path.
//
// Every record of a log (quad, 4 cells x 5 fields of FP32 values) counts round(q[f] * v)
// units at each of its cells, q[f] the Loc grid's fixed-point scale of field f, the rounding
// done per record.  The Loc grids of the FP32 path are therefore always exactly the sum over
// the particles of their current orbits' records, whatever the order of the adds: the initial
// pass adds every orbit's records, and taking a trial adds its orbit's records and takes away
// those of the orbit it replaces (the reference keeps Loc += response(new) - response(old),
// main.c:503-513; here both responses are the logged FP32 ones).
//
// Few global atomics: the records are first bucketed by region of GS x GS cells (32 x 32, or
// the whole grid below 32 x 32) -- k_commit_count counts the entries of each region (LDS
// counts per particle, one global atomic per particle and region), the host (or k_commit_layout,
// where the commit is captured in a graph) turns the counts into bucket offsets and chunks,
// k_commit_scatter writes (record, sign) entries into the
// buckets -- and k_commit_regions sums a chunk of one region's entries in LDS (int64 atomics)
// and adds each nonzero cell once to the destination grid.

#include <hip/hip_runtime.h>

#include "galic_fast.h"

namespace galic {

namespace {

// Regions (index gz * nreg + gr, nreg regions per side) reached by the 2x2 cells of quad.
template <typename Visit>
__device__ __forceinline__ void region_visit(int quad, int L, int gsh, Visit visit) {
  const int iz = quad >> L, iR = quad & ((1 << L) - 1), nreg = 1 << (L - gsh);
  const int z0 = iz >> gsh, z1 = (iz + 1) >> gsh, r0 = iR >> gsh, r1 = (iR + 1) >> gsh;
  visit(z0 * nreg + r0);
  if(r1 != r0) visit(z0 * nreg + r1);
  if(z1 != z0) {
    visit(z1 * nreg + r0);
    if(r1 != r0) visit(z1 * nreg + r1);
  }
}

// The logs item b of a commit launch adds (sign 0) or takes away (sign 1): with accept, the
// taken trial's orbit (+) and the particle's current orbit (-), nothing if no trial was taken;
// without, orbit b (+).  Returns the number of logs (0..2) and the particle type.
__device__ __forceinline__ int commit_logs(const CommitArgs &A, int b, size_t (&orb)[2], unsigned int (&sign)[2],
                                           int &t) {
  if(!A.accept) {
    orb[0] = (size_t)A.orbit_base + b;
    sign[0] = 0;
    t = A.otype[b];
    return 1;
  }
  const int ti = A.trial_base + b;
  const int best = A.accept[ti] - 1;
  t = A.ptype[A.tidx[ti]];
  if(best < 0) return 0;
  const size_t o0 = (size_t)(A.slot_pos ? A.slot_pos[b] : b) * (A.K + 1);
  orb[0] = o0 + 1 + best;
  sign[0] = 0;
  orb[1] = o0;
  sign[1] = 1;
  return 2;
}

// Calls visit(record index, sign) for every record of the item's logs (all threads).
template <typename Visit>
__device__ __forceinline__ void commit_records(const CommitArgs &A, int nlog, const size_t (&orb)[2],
                                               const unsigned int (&sign)[2], Visit visit) {
  for(int l = 0; l < nlog; l++) {
    const int n = max(A.log_n[orb[l]], 0);
    const int *tb = A.log_blocks + orb[l] * A.log_M;
    for(int r = threadIdx.x; r < n; r += blockDim.x) {
      const unsigned int rec = (unsigned int)tb[r / LOG_B] * LOG_B + (unsigned int)(r & (LOG_B - 1));
      visit(rec, sign[l]);
    }
  }
}

}  // namespace

// Entries per (type, region): one workgroup per item, counts in LDS, then one global atomic
// per region the item reaches.
__global__ void __launch_bounds__(256) k_commit_count(CommitArgs A) {
  extern __shared__ unsigned int cnt_s[];
  const int nreg = 1 << (A.L - A.gsh), ng = nreg * nreg;
  for(int b = blockIdx.x; b < A.n; b += gridDim.x) {
    size_t orb[2];
    unsigned int sign[2];
    int t;
    const int nlog = commit_logs(A, b, orb, sign, t);
    if(nlog == 0) continue;
    for(int g = threadIdx.x; g < ng; g += blockDim.x) cnt_s[g] = 0;
    __syncthreads();
    commit_records(A, nlog, orb, sign, [&](unsigned int rec, unsigned int) {
      region_visit(A.log_q[rec], A.L, A.gsh, [&](int g) { atomicAdd(&cnt_s[g], 1u); });
    });
    __syncthreads();
    for(int g = threadIdx.x; g < ng; g += blockDim.x)
      if(cnt_s[g]) atomicAdd(&A.gcnt[(size_t)t * ng + g], cnt_s[g]);
    __syncthreads();
  }
}

// The entries into the buckets: per item, LDS counts, one global atomic per region to reserve
// room in its bucket (gstart + cursor), then the entries at LDS cursors.
__global__ void __launch_bounds__(256) k_commit_scatter(CommitArgs A) {
  extern __shared__ unsigned int cnt_s[];
  if(A.gerr && *A.gerr) return;
  const int nreg = 1 << (A.L - A.gsh), ng = nreg * nreg;
  for(int b = blockIdx.x; b < A.n; b += gridDim.x) {
    size_t orb[2];
    unsigned int sign[2];
    int t;
    const int nlog = commit_logs(A, b, orb, sign, t);
    if(nlog == 0) continue;
    for(int g = threadIdx.x; g < ng; g += blockDim.x) cnt_s[g] = 0;
    __syncthreads();
    commit_records(A, nlog, orb, sign, [&](unsigned int rec, unsigned int) {
      region_visit(A.log_q[rec], A.L, A.gsh, [&](int g) { atomicAdd(&cnt_s[g], 1u); });
    });
    __syncthreads();
    for(int g = threadIdx.x; g < ng; g += blockDim.x)
      if(cnt_s[g]) {
        const size_t k = (size_t)t * ng + g;
        cnt_s[g] = A.gstart[k] + atomicAdd(&A.gcur[k], cnt_s[g]);
      }
    __syncthreads();
    commit_records(A, nlog, orb, sign, [&](unsigned int rec, unsigned int s) {
      region_visit(A.log_q[rec], A.L, A.gsh,
                   [&](int g) { A.gbuf[atomicAdd(&cnt_s[g], 1u)] = make_uint2(rec, s); });
    });
    __syncthreads();
  }
}

// One workgroup per chunk (type, region, entries [lo, hi) of its bucket): each entry's cells
// inside the region, round(q[f] * v) with the entry's sign, into five int64 LDS grids of the
// region; then every nonzero cell once into dst[type][f][cell] (a global atomic).
// (nchunk: the number of chunks when the grid was sized for at most that many; null when the
// grid is exactly one workgroup per chunk)
__global__ void __launch_bounds__(256) k_commit_regions(CommitArgs A, const int4 *chunks, long long *dst,
                                                        const unsigned int *nchunk) {
  extern __shared__ long long reg_s[];  // [5][GS * GS]
  if(nchunk && blockIdx.x >= *nchunk) return;
  const int gs = 1 << A.gsh, nreg = 1 << (A.L - A.gsh), nb = 1 << A.L, ngrid = nb * nb;
  const int4 ch = chunks[blockIdx.x];  // (type * nreg^2 + region, lo, hi, 0)
  const int t = ch.x / (nreg * nreg), g = ch.x % (nreg * nreg);
  const int gz0 = (g / nreg) * gs, gr0 = (g % nreg) * gs;
  const int cells = gs * gs;
  for(int q = threadIdx.x; q < kFields * cells; q += blockDim.x) reg_s[q] = 0;
  __syncthreads();
  for(int e = ch.y + threadIdx.x; e < ch.z; e += blockDim.x) {
    const uint2 en = A.gbuf[e];
    const unsigned int rec = en.x;
    const int quad = A.log_q[rec], blk = (int)(rec / LOG_B), jj = (int)(rec % LOG_B);
    const int lz = (quad >> A.L) - gz0, lr = (quad & (nb - 1)) - gr0;
#pragma unroll
    for(int f = 0; f < kFields; f++) {
      const float4 x = A.log_v[((size_t)blk * LOG_B + jj) * kFields + f];
      const double q = en.y ? -A.q[f] : A.q[f];
#pragma unroll
      for(int c = 0; c < 4; c++) {
        const int cz = lz + (c >> 1), cr = lr + (c & 1);
        const float v = c == 0 ? x.x : c == 1 ? x.y : c == 2 ? x.z : x.w;
        if((unsigned)cz < (unsigned)gs && (unsigned)cr < (unsigned)gs && v != 0.f)
          atomicAdd((unsigned long long *)&reg_s[f * cells + cz * gs + cr],
                    (unsigned long long)__double2ll_rn(q * (double)v));
      }
    }
  }
  __syncthreads();
  long long *d = dst + (size_t)t * kFields * ngrid;
  for(int q = threadIdx.x; q < kFields * cells; q += blockDim.x) {
    const long long v = reg_s[q];
    if(v) {
      const int f = q / cells, cz = (q % cells) / gs, cr = q % gs;
      atomicAdd((unsigned long long *)&d[(size_t)f * ngrid + (size_t)(gz0 + cz) * nb + gr0 + cr],
                (unsigned long long)v);
    }
  }
}

// The bucket layout of k_commit_regions on the device (one workgroup of 1024): gstart[k] the
// first entry of bucket k (the buckets in order), and chunks of at most `chunk` entries of one
// bucket each, nchunk of them.  If the entries or chunks exceed the room (cap_entries,
// cap_chunks), nothing is laid out: nchunk = 0 and *err = 1.
__global__ void __launch_bounds__(1024) k_commit_layout(const unsigned int *gcnt, int nk, int chunk,
                                                        unsigned int *gstart, int4 *chunks, unsigned int *nchunk,
                                                        unsigned int cap_entries, unsigned int cap_chunks,
                                                        unsigned int *err) {
  __shared__ unsigned long long s_e[1024];
  __shared__ unsigned int s_c[1024];
  const int tid = threadIdx.x, per = (nk + 1023) / 1024;
  const int lo = min(nk, tid * per), hi = min(nk, lo + per);
  unsigned long long e = 0;
  unsigned int c = 0;
  for(int k = lo; k < hi; k++) {
    e += gcnt[k];
    c += (gcnt[k] + chunk - 1) / chunk;
  }
  s_e[tid] = e;
  s_c[tid] = c;
  __syncthreads();
  for(int off = 1; off < 1024; off <<= 1) {
    const unsigned long long ve = tid >= off ? s_e[tid - off] : 0;
    const unsigned int vc = tid >= off ? s_c[tid - off] : 0;
    __syncthreads();
    s_e[tid] += ve;
    s_c[tid] += vc;
    __syncthreads();
  }
  const unsigned long long tot_e = s_e[1023];
  const unsigned int tot_c = s_c[1023];
  if(tot_e > cap_entries || tot_c > cap_chunks) {
    if(tid == 0) {
      *nchunk = 0;
      *err = 1;
    }
    return;
  }
  unsigned int e0 = (unsigned int)(s_e[tid] - e), c0 = s_c[tid] - c;
  for(int k = lo; k < hi; k++) {
    const unsigned int n = gcnt[k];
    gstart[k] = e0;
    for(unsigned int l = 0; l < n; l += chunk)
      chunks[c0++] = make_int4(k, (int)(e0 + l), (int)(e0 + min(n, l + chunk)), 0);
    e0 += n;
  }
  if(tid == 0) *nchunk = tot_c;
}

// Room a commit of trials [0, n) of a group needs, per rep: a trial adds at most its largest
// trial orbit's records and takes away its current orbit's, each record in at most 4 region
// buckets.  Trial b is trial first + b, of rep (first + b) / rep_size - first / rep_size.
__global__ void k_commit_bound(int n, int K, int first, int rep_size, const int *slot_pos, const int *log_n,
                               unsigned long long *rep_bound) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if(b >= n) return;
  const size_t o0 = (size_t)slot_pos[b] * (K + 1);
  int mx = 0;
  for(int k = 1; k <= K; k++) mx = max(mx, log_n[o0 + k]);
  const unsigned long long e = 4ull * ((unsigned long long)mx + (unsigned long long)max(log_n[o0], 0));
  atomicAdd(&rep_bound[(first + b) / rep_size - first / rep_size], e);
}

}  // namespace galic
