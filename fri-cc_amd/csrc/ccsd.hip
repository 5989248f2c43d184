// Fused end of the amplitude update: symmetrisation, rank-one t1 terms,
// denominators, energy and amplitude-change norm in one pass over T2.
//
// The reference builds t2new from many full-size temporaries
// (tmp + tmp.transpose(1,0,3,2) four times, an explicit eijab array,
// fricc/friccsd.py:178-284) and then computes the energy and the norm of the
// change in separate passes (fricc/friccsd.py:418-427).  Here every
// symmetrised contribution is accumulated into one buffer S, every direct one
// into the output buffer itself, and this kernel finishes the element:
//
//   t2new[ijab] = ( ovovT[ijab] + acc[ijab] + S[ijab] + S[jiba]
//                   + t1[i,a] t1[j,b] Wsum[a,b] ) / (eia[i,a] + eia[j,b])
//
// and reduces, in FP64,
//   E2    = sum_ijab (t2new[ijab] + t1new[i,a] t1new[j,b]) * Lovov[ijab]
//   dnorm = sum over ia >= jb of (t2new[ijab] - t2old[ijab])^2
// (the packed lower triangle of t2 as an (ia, jb) matrix,
// ccsd.amplitudes_to_vector).  Partial sums are written per block and added
// in a fixed order by a second kernel, so the result is deterministic.
#include "fricc_internal.h"

namespace fricc {

constexpr int FT = 64;  // (a, b) tile edge
constexpr int FR = 4;   // rows of a tile per pass of the 256 threads (64 columns each)

template <typename TS>
__global__ __launch_bounds__(256) void k_t2_finalize(int o, int v, const double* __restrict__ ovovT,
                                                      double* __restrict__ acc, const TS* __restrict__ S,
                                                      const double* __restrict__ t1, const double* __restrict__ wsum,
                                                      const double* __restrict__ eia,
                                                      const double* __restrict__ t2old,
                                                      const double* __restrict__ lovov,
                                                      const double* __restrict__ t1new,
                                                      double* __restrict__ partial) {
  __shared__ double st[FT][FT + 1];
  __shared__ double red[2][256 / WAVE];
  // one block per (i, j, 64-row band of a)
  const int nat = (v + FT - 1) / FT;
  const int pair = blockIdx.x / nat, a0 = (blockIdx.x % nat) * FT;
  const int i = pair / o, j = pair % o;
  const int64_t vv = (int64_t)v * v;
  const int64_t pij = (int64_t)(i * o + j) * vv, pji = (int64_t)(j * o + i) * vv;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
  double e2 = 0.0, dn = 0.0;
  for (int b0 = 0; b0 < v; b0 += FT) {
    // S[j,i,b,a] for b in [b0,b0+64), a in [a0,a0+64): a is contiguous
#pragma unroll 4
    for (int r = ty; r < FT; r += FR) {
      const int b = b0 + r, a = a0 + tx;
      st[r][tx] = (b < v && a < v) ? (double)S[pji + (int64_t)b * v + a] : 0.0;
    }
    __syncthreads();
    // 16 rows per thread, unrolled so that their loads are in flight together
#pragma unroll 4
    for (int r = ty; r < FT; r += FR) {
      const int a = a0 + r, b = b0 + tx;
      if (a < v && b < v) {
        const int64_t e = pij + (int64_t)a * v + b;
        double val = ovovT[e] + acc[e] + (double)S[e] + st[tx][r];
        if (wsum) val += t1[i * v + a] * t1[j * v + b] * wsum[a * v + b];
        const double t2n = val / (eia[i * v + a] + eia[j * v + b]);
        acc[e] = t2n;
        e2 += (t2n + t1new[i * v + a] * t1new[j * v + b]) * lovov[e];
        if (i * v + a >= j * v + b) {
          const double d = t2n - t2old[e];
          dn += d * d;
        }
      }
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  for (int d = WAVE / 2; d > 0; d >>= 1) {
    e2 += __shfl_down(e2, d, WAVE);
    dn += __shfl_down(dn, d, WAVE);
  }
  if (lane == 0) {
    red[0][w] = e2;
    red[1][w] = dn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < 256 / WAVE; k++) {
      a += red[0][k];
      b += red[1][k];
    }
    partial[2 * blockIdx.x] = a;
    partial[2 * blockIdx.x + 1] = b;
  }
}

// One block of 256 threads: thread t adds the partials t, t+256, ... in order,
// then a fixed-shape tree adds the 256 results, so the sum does not depend on
// timing.
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partial, int n,
                                                      double* __restrict__ out) {
  __shared__ double red[2][256];
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < n; k += 256) {
    a += partial[2 * k];
    b += partial[2 * k + 1];
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = red[0][0];
    out[1] = red[1][0];
  }
}

int t2_finalize(int o, int v, const double* ovovT, double* acc, const void* S, int s_f32, const double* t1,
                const double* wsum, const double* eia, const double* t2old, const double* lovov, const double* t1new,
                double* sums, hipStream_t s) {
  FRICC_CHECK(o > 0 && v > 0, "t2_finalize: bad dimensions");
  void* p;
  const int nb = o * o * ((v + FT - 1) / FT);
  FRICC_TRY(ws_get(40, (size_t)nb * 2 * 8, &p));
  double* partial = (double*)p;
  if (s_f32)
    k_t2_finalize<float><<<nb, 256, 0, s>>>(o, v, ovovT, acc, (const float*)S, t1, wsum, eia, t2old, lovov, t1new,
                                            partial);
  else
    k_t2_finalize<double><<<nb, 256, 0, s>>>(o, v, ovovT, acc, (const double*)S, t1, wsum, eia, t2old, lovov, t1new,
                                             partial);
  k_sum_partials<<<1, 256, 0, s>>>(partial, nb, sums);
  FRICC_HIP(hipGetLastError());
  return OK;
}

}  // namespace fricc

extern "C" {
// Fused end of the amplitude update (see the top of this file).  All arrays
// are device C-order: ovovT, acc (in/out: direct terms in, t2new out), t2old,
// lovov [o][o][v][v] FP64; S [o][o][v][v] FP64 or FP32 (s_precision);
// t1, eia, t1new [o][v]; wsum [v][v] or NULL.  sums (device, 2 doubles)
// receives {E2, dnorm}.  No counterpart in the reference.
int fricc_t2_finalize(int o, int v, const double* ovovT, double* acc, const void* S, int s_precision,
                      const double* t1, const double* wsum, const double* eia, const double* t2old,
                      const double* lovov, const double* t1new, double* sums, void* stream) {
  FRICC_CHECK(s_precision == 0 || s_precision == 1, "t2_finalize: S must be FP64 or FP32");
  return fricc::t2_finalize(o, v, ovovT, acc, S, s_precision, t1, wsum, eia, t2old, lovov, t1new, sums,
                            (hipStream_t)stream);
}
}
