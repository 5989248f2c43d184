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
// ccsd.amplitudes_to_vector).  Lovov[ijab] = 2 ovovT[ijab] - ovovT[ijba] is
// either read or, when the caller passes no Lovov, formed in the kernel from
// the ovovT tiles it already holds.  Partial sums are written per block and
// added in a fixed order by a second kernel, so the result is deterministic.
#include "fricc_internal.h"

namespace fricc {

constexpr int QE = 32;  // edge of one (a, b) tile of a quad

// k -> (lo, hi) with lo <= hi < n, k = lo*n - lo*(lo-1)/2 + (hi - lo)
__device__ __forceinline__ void tri_decode(int k, int n, int& lo, int& hi) {
  int a = 0;
  while (k >= n - a) {
    k -= n - a;
    a++;
  }
  lo = a;
  hi = a + k;
}

template <int VEC, typename T>
__device__ __forceinline__ void ldv(const T* __restrict__ p, T (&d)[VEC]) {
  if constexpr (VEC == 2) {
    if constexpr (sizeof(T) == 8) {
      const double2 t = *reinterpret_cast<const double2*>(p);
      d[0] = t.x;
      d[1] = t.y;
    } else {
      const float2 t = *reinterpret_cast<const float2*>(p);
      d[0] = t.x;
      d[1] = t.y;
    }
  } else {
    d[0] = p[0];
  }
}

// One block per quad: an unordered occupied pair {i <= j} and an unordered
// pair of 32-wide virtual tiles {A <= B}.  Tile q of the quad is the pair
// (i,j) for q < 2 and (j,i) otherwise, with rows in A and columns in B for
// even q and the reverse for odd q.  Every element's partners then lie in the
// same quad: the exchange partner ovovT[P,Q,s,f] of (P,Q,f,s) in tile q^1 and
// the symmetrisation partner S[Q,P,s,f] in tile 3-q, both at the transposed
// position, so each input is read from HBM once, row by row, and the
// transposes go through LDS.  With lovov NULL (READL false) the exchange
// integral 2 ovovT[ijab] - ovovT[ijba] is formed from the two ovovT tiles
// instead of being read.  On the diagonal (i == j or A == B) some tiles
// coincide; only the distinct ones are read, written and summed.
template <typename TS, int VEC, bool READL>
__global__ __launch_bounds__(256) void k_t2_finalize_quad(int o, int v, const double* __restrict__ ovovT,
                                                           double* __restrict__ acc, const TS* __restrict__ S,
                                                           const double* __restrict__ t1,
                                                           const double* __restrict__ wsum,
                                                           const double* __restrict__ eia,
                                                           const double* __restrict__ t2old,
                                                           const double* __restrict__ lovov,
                                                           const double* __restrict__ t1new,
                                                           float* __restrict__ out32,
                                                           double* __restrict__ partial) {
  __shared__ double xs[4][QE][QE + 1];
  __shared__ TS ss[4][QE][QE + 1];
  __shared__ double red[2][256 / WAVE];
  constexpr int LPR = QE / VEC, RPP = 256 / LPR, NP = QE / RPP;  // lanes per row, rows per pass, passes
  int i, j, ta, tb;
  tri_decode(blockIdx.x, o, i, j);
  tri_decode(blockIdx.y, (v + QE - 1) / QE, ta, tb);
  const int A0 = ta * QE, B0 = tb * QE;
  // tile q duplicates tile q & dmask
  int dmask = 3;
  if (i == j) dmask &= 1;
  if (ta == tb) dmask &= 2;
  const int64_t vv = (int64_t)v * v;
  const int cl = (threadIdx.x % LPR) * VEC, rl = threadIdx.x / LPR;
  double x[4][NP][VEC], ac[4][NP][VEC], to[4][NP][VEC], lv[4][NP][VEC];
  TS sv[4][NP][VEC];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    if ((q & dmask) != q) continue;
    const int P = (q >> 1) ? j : i, Q = (q >> 1) ? i : j;
    const int R0 = (q & 1) ? B0 : A0, C0 = (q & 1) ? A0 : B0;
    const int64_t base = (int64_t)(P * o + Q) * vv;
#pragma unroll
    for (int p = 0; p < NP; p++) {
      const int r = rl + p * RPP, f = R0 + r, s = C0 + cl;
      if (f < v && s < v) {
        const int64_t e = base + (int64_t)f * v + s;
        ldv<VEC>(ovovT + e, x[q][p]);
        ldv<VEC>(S + e, sv[q][p]);
        ldv<VEC>(acc + e, ac[q][p]);
        ldv<VEC>(t2old + e, to[q][p]);
        if constexpr (READL) ldv<VEC>(lovov + e, lv[q][p]);
      } else {
#pragma unroll
        for (int k = 0; k < VEC; k++) {
          x[q][p][k] = 0.0;
          sv[q][p][k] = TS(0);
        }
      }
#pragma unroll
      for (int k = 0; k < VEC; k++) {
        xs[q][r][cl + k] = x[q][p][k];
        ss[q][r][cl + k] = sv[q][p][k];
      }
    }
  }
  __syncthreads();
  double e2 = 0.0, dn = 0.0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    if ((q & dmask) != q) continue;
    const int P = (q >> 1) ? j : i, Q = (q >> 1) ? i : j;
    const int R0 = (q & 1) ? B0 : A0, C0 = (q & 1) ? A0 : B0;
    const int qx = (q ^ 1) & dmask, qs = (3 - q) & dmask;
    const int64_t base = (int64_t)(P * o + Q) * vv;
#pragma unroll
    for (int p = 0; p < NP; p++) {
      const int r = rl + p * RPP, f = R0 + r, s0 = C0 + cl;
      if (f >= v || s0 >= v) continue;
      const double ef = eia[P * v + f], tf = t1[P * v + f], nf = t1new[P * v + f];
      double out[VEC];
#pragma unroll
      for (int k = 0; k < VEC; k++) {
        const int c = cl + k, s = s0 + k;
        double val = x[q][p][k] + ac[q][p][k] + (double)sv[q][p][k] + (double)ss[qs][c][r];
        if (wsum) val += tf * t1[Q * v + s] * wsum[(int64_t)f * v + s];
        const double t2n = val / (ef + eia[Q * v + s]);
        out[k] = t2n;
        const double L = READL ? lv[q][p][k] : 2.0 * x[q][p][k] - xs[qx][c][r];
        e2 += (t2n + nf * t1new[Q * v + s]) * L;
        if (P * v + f >= Q * v + s) {
          const double d = t2n - to[q][p][k];
          dn += d * d;
        }
      }
      double* dst = acc + base + (int64_t)f * v + s0;
      if constexpr (VEC == 2)
        *reinterpret_cast<double2*>(dst) = make_double2(out[0], out[1]);
      else
        dst[0] = out[0];
      if (out32) {
        float* d32 = out32 + base + (int64_t)f * v + s0;
        if constexpr (VEC == 2)
          *reinterpret_cast<float2*>(d32) = make_float2((float)out[0], (float)out[1]);
        else
          d32[0] = (float)out[0];
      }
    }
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
    partial[2 * ((int64_t)blockIdx.y * gridDim.x + blockIdx.x)] = a;
    partial[2 * ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) + 1] = b;
  }
}

// One block of SPT threads: thread t adds the partials t, t+SPT, ... in
// order (eight loads in flight at a time, added in that same order), then a
// fixed-shape tree adds the SPT results, so the sum does not depend on timing.
constexpr int SPT = 1024;
__global__ __launch_bounds__(SPT) void k_sum_partials(const double* __restrict__ partial, int n,
                                                      double* __restrict__ out) {
  __shared__ double red[2][SPT];
  double a = 0.0, b = 0.0;
  for (int k0 = threadIdx.x; k0 < n; k0 += 8 * SPT) {
    double2 u[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = k0 + q * SPT;
      u[q] = k < n ? *reinterpret_cast<const double2*>(partial + 2 * k) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
      if (k0 + q * SPT < n) {
        a += u[q].x;
        b += u[q].y;
      }
    }
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  __syncthreads();
  for (int h = SPT / 2; h > 0; h >>= 1) {
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

template <typename TS, int VEC>
static void finalize_launch(dim3 grid, int o, int v, const double* ovovT, double* acc, const void* S,
                            const double* t1, const double* wsum, const double* eia, const double* t2old,
                            const double* lovov, const double* t1new, float* out32, double* partial,
                            hipStream_t s) {
  if (lovov)
    k_t2_finalize_quad<TS, VEC, true><<<grid, 256, 0, s>>>(o, v, ovovT, acc, (const TS*)S, t1, wsum, eia, t2old,
                                                           lovov, t1new, out32, partial);
  else
    k_t2_finalize_quad<TS, VEC, false><<<grid, 256, 0, s>>>(o, v, ovovT, acc, (const TS*)S, t1, wsum, eia, t2old,
                                                            lovov, t1new, out32, partial);
}

int t2_finalize(int o, int v, const double* ovovT, double* acc, const void* S, int s_f32, const double* t1,
                const double* wsum, const double* eia, const double* t2old, const double* lovov, const double* t1new,
                float* out32, double* sums, hipStream_t s) {
  FRICC_CHECK(o > 0 && v > 0, "t2_finalize: bad dimensions");
  const int nt = (v + QE - 1) / QE;
  const int64_t npair = (int64_t)o * (o + 1) / 2, ntile = (int64_t)nt * (nt + 1) / 2;
  FRICC_CHECK(npair < (1LL << 31) && ntile <= 65535 && npair * ntile < (1LL << 31),
              "t2_finalize: dimensions too large");
  const int64_t nb = npair * ntile;
  void* p;
  FRICC_TRY(ws_get(40, (size_t)nb * 2 * 8, &p));
  double* partial = (double*)p;
  // 16-byte accesses need an even v and 16-byte aligned operands (8 for an FP32 S)
  auto al = [](const void* q, uintptr_t n) { return q == nullptr || ((uintptr_t)q % n) == 0; };
  const bool vec = v % 2 == 0 && al(ovovT, 16) && al(acc, 16) && al(t2old, 16) && al(lovov, 16) &&
                   al(S, s_f32 ? 8 : 16) && al(out32, 8);
  const dim3 grid((unsigned)npair, (unsigned)ntile);
  if (s_f32) {
    if (vec)
      finalize_launch<float, 2>(grid, o, v, ovovT, acc, S, t1, wsum, eia, t2old, lovov, t1new, out32, partial, s);
    else
      finalize_launch<float, 1>(grid, o, v, ovovT, acc, S, t1, wsum, eia, t2old, lovov, t1new, out32, partial, s);
  } else {
    if (vec)
      finalize_launch<double, 2>(grid, o, v, ovovT, acc, S, t1, wsum, eia, t2old, lovov, t1new, out32, partial, s);
    else
      finalize_launch<double, 1>(grid, o, v, ovovT, acc, S, t1, wsum, eia, t2old, lovov, t1new, out32, partial, s);
  }
  k_sum_partials<<<1, SPT, 0, s>>>(partial, (int)nb, sums);
  FRICC_HIP(hipGetLastError());
  return OK;
}

}  // namespace fricc

extern "C" {
// Fused end of the amplitude update (see the top of this file).  All arrays
// are device C-order: ovovT, acc (in/out: direct terms in, t2new out), t2old,
// lovov [o][o][v][v] FP64 or NULL (then 2 ovovT[ijab] - ovovT[ijba] is used);
// S [o][o][v][v] FP64 or FP32 (s_precision); t1, eia, t1new [o][v]; wsum
// [v][v] or NULL; t2new_f32 [o][o][v][v] FP32 or NULL (also receives t2new,
// rounded to FP32).  sums (device, 2 doubles)
// receives {E2, dnorm}.  No counterpart in the reference.
int fricc_t2_finalize(int o, int v, const double* ovovT, double* acc, const void* S, int s_precision,
                      const double* t1, const double* wsum, const double* eia, const double* t2old,
                      const double* lovov, const double* t1new, float* t2new_f32, double* sums, void* stream) {
  FRICC_CHECK(s_precision == 0 || s_precision == 1, "t2_finalize: S must be FP64 or FP32");
  return fricc::t2_finalize(o, v, ovovT, acc, S, s_precision, t1, wsum, eia, t2old, lovov, t1new, t2new_f32, sums,
                            (hipStream_t)stream);
}
}
