// Two-electron integrals from density-fitting factors on the GPU:
//   (pq|rs) = sum_Q L[Q,p,q] L[Q,r,s]
// evaluated as a sequential fused-multiply-add chain over Q = 0..naux-1
// starting from 0.  Since fma(a,b,c) == fma(b,a,c) and L is symmetric in
// (p,q), every member of an 8-fold orbit gets bitwise the same value, and the
// CPU oracle (same chain) reproduces every integral bit for bit.
//
// The kernel is GEMM-tiled over pair space: rows are (p,q) pairs, columns
// (r,s) pairs; each thread keeps a 4x4 register tile and walks Q in order, so
// the per-element chain is exactly the sequential one.  Output positions are
// given by four strides, so any block can be written directly in the layout a
// consumer wants (e.g. the ladder operand W[c,d,a,b] = (ac|bd)).
#include "fricc_internal.h"

namespace fricc {

constexpr int E_T = 64;   // tile edge (pairs)
constexpr int E_QC = 16;  // Q chunk staged in LDS

struct EriGeom {
  int64_t nmo, naux;
  int64_t p0, q0, r0, s0;
  int64_t nq, ns;          // inner extents of the pair indices
  int64_t nrow, ncol;      // number of (p,q) and (r,s) pairs
  int64_t sp, sq, sr, ss;  // output strides
};

__global__ __launch_bounds__(256) void k_eri(const double* __restrict__ L, EriGeom g, double* __restrict__ out) {
  __shared__ double la[E_QC][E_T];
  __shared__ double lb[E_QC][E_T];
  const int64_t rt = (int64_t)blockIdx.y * E_T, ct = (int64_t)blockIdx.x * E_T;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int64_t n2 = g.nmo * g.nmo;
  // each thread loads 4 (Q, pair) entries of each operand per chunk
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
  const int lq = threadIdx.x >> 6;  // 0..3
  const int lp = threadIdx.x & 63;
  int64_t offa = -1, offb = -1;
  {
    int64_t pr = rt + lp;
    if (pr < g.nrow) offa = (g.p0 + pr / g.nq) * g.nmo + (g.q0 + pr % g.nq);
    int64_t pc = ct + lp;
    if (pc < g.ncol) offb = (g.r0 + pc / g.ns) * g.nmo + (g.s0 + pc % g.ns);
  }
  for (int64_t Q0 = 0; Q0 < g.naux; Q0 += E_QC) {
#pragma unroll
    for (int k = 0; k < E_QC / 4; k++) {
      const int q = lq + 4 * k;
      const int64_t Q = Q0 + q;
      la[q][lp] = (Q < g.naux && offa >= 0) ? L[Q * n2 + offa] : 0.0;
      lb[q][lp] = (Q < g.naux && offb >= 0) ? L[Q * n2 + offb] : 0.0;
    }
    __syncthreads();
    const int qn = (g.naux - Q0) < E_QC ? (int)(g.naux - Q0) : E_QC;
    for (int q = 0; q < qn; q++) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = la[q][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = lb[q][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int64_t pr = rt + ty + 16 * i;
    if (pr >= g.nrow) continue;
    const int64_t p = pr / g.nq, q = pr % g.nq;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int64_t pc = ct + tx + 16 * j;
      if (pc >= g.ncol) continue;
      const int64_t r = pc / g.ns, s = pc % g.ns;
      out[p * g.sp + q * g.sq + r * g.sr + s * g.ss] = acc[i][j];
    }
  }
}

int eri_build(const double* L, int64_t naux, int64_t nmo, const int64_t rng[8], const int64_t st[4], double* out,
              hipStream_t s) {
  EriGeom g;
  g.nmo = nmo;
  g.naux = naux;
  g.p0 = rng[0];
  g.q0 = rng[2];
  g.r0 = rng[4];
  g.s0 = rng[6];
  const int64_t np = rng[1] - rng[0], nq = rng[3] - rng[2], nr = rng[5] - rng[4], ns = rng[7] - rng[6];
  FRICC_CHECK(np > 0 && nq > 0 && nr > 0 && ns > 0, "eri_build: empty range");
  FRICC_CHECK(rng[1] <= nmo && rng[3] <= nmo && rng[5] <= nmo && rng[7] <= nmo && rng[0] >= 0 && rng[2] >= 0 &&
                  rng[4] >= 0 && rng[6] >= 0,
              "eri_build: range outside the orbital space");
  g.nq = nq;
  g.ns = ns;
  g.nrow = np * nq;
  g.ncol = nr * ns;
  g.sp = st[0];
  g.sq = st[1];
  g.sr = st[2];
  g.ss = st[3];
  const int64_t gy = ceil_div(g.nrow, E_T), gx = ceil_div(g.ncol, E_T);
  FRICC_CHECK(gy < 65536, "eri_build: too many (p,q) pairs for one launch");
  k_eri<<<dim3((unsigned)gx, (unsigned)gy), 256, 0, s>>>(L, g, out);
  FRICC_HIP(hipGetLastError());
  return OK;
}

}  // namespace fricc

extern "C" {
// (pq|rs) for p in [rng0,rng1), q in [rng2,rng3), r in [rng4,rng5), s in
// [rng6,rng7) from device L[naux][nmo][nmo]; element written at
// out[(p-rng0)*st0 + (q-rng2)*st1 + (r-rng4)*st2 + (s-rng6)*st3].
// Synthetic-integral provider; no counterpart in the reference (which takes
// PySCF's ao2mo integrals, fricc/friccsd.py:503-504).
int fricc_eri_build(const double* L, int64_t naux, int64_t nmo, const int64_t rng[8], const int64_t strides[4],
                    double* out, void* stream) {
  return fricc::eri_build(L, naux, nmo, rng, strides, out, (hipStream_t)stream);
}
}
