// Gathered dense-tile contraction: OUT[x, y] += alpha * sum_k A[rowlist[k], x] * B[k, y].
//
// Replaces the 11 loop nests of src/sparse_tensor.cpp:11-346.  The reference
// loops over output blocks and rescans all m entries per block; here the
// sparse operand is a View (entries gathered per output-column tile into a
// dense [rows x TY] block) and the dense operand A is read row-by-row for the
// gathered rows only, so each tile is one MFMA GEMM with a gathered K axis.
//
// Tiling (f64 / f32, v_mfma_*_16x16x4): a 256-thread workgroup owns a
// 128 (x) x 64 (y) output tile; wave w owns x rows [32w, 32w+32) and all 64
// columns: 2 x 4 MFMA 16x16 blocks.  Per K step (4 gathered rows) a lane
// loads one A value per 16-row x block (16 consecutive x of one gathered row:
// 128 B segments) and one B value per 16-column block.
#include "fricc_internal.h"

namespace fricc {

typedef __attribute__((ext_vector_type(4))) double d4;
typedef __attribute__((ext_vector_type(4))) float f4;

struct OutDev {
  void* out;
  int of32;
  uint32_t Xl, Yl;
  int64_t oxh, oxl, oyh, oyl;
};

// out[off] += alpha * v in the output's element type
__device__ inline void out_add(const OutDev& om, int64_t off, double alpha, double v) {
  if (om.of32)
    ((float*)om.out)[off] += (float)(alpha * v);
  else
    ((double*)om.out)[off] += alpha * v;
}

template <typename T>
struct Acc;
template <>
struct Acc<double> {
  typedef d4 type;
  static __device__ inline d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // D register i of lane group g holds row g + 4 i (measured on gfx950)
  static __device__ inline int row(int g, int i) { return g + 4 * i; }
};
template <>
struct Acc<float> {
  typedef f4 type;
  static __device__ inline f4 mfma(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // D register i of lane group g holds row 4 g + i
  static __device__ inline int row(int g, int i) { return 4 * g + i; }
};

constexpr int GT_X = 256;
constexpr int GT_Y = 64;

// Blocks are numbered so that consecutive ids on one XCD (ids are dealt
// round-robin over the 8 XCDs) walk the x tiles of one column tile, which
// share the same B block and gathered A rows in that XCD's L2.
// order 0: x tiles fastest (B block reused from L2); order 1: column tiles
// fastest (the gathered A rows of one x range reused across column tiles).
__device__ inline void xcd_remap(int64_t nbx, int64_t nty, int order, int64_t& bx, int64_t& ty) {
  const int64_t nb = nbx * nty;
  int64_t id = blockIdx.x;
  const int64_t per = nb / 8;
  if (per * 8 == nb) id = (id % 8) * per + id / 8;
  if (order == 0) {
    ty = id / nbx;
    bx = id % nbx;
  } else {
    bx = id / nty;
    ty = id % nty;
  }
}

// Each wave owns a 64 (x) x 64 (y) tile = 4 x 4 MFMA blocks; the 4 waves of a
// workgroup split 256 consecutive x.  K advances in chunks of 16 gathered rows
// (4 MFMA K steps); the loads of chunk c+1 are issued into a second register
// set before the MFMAs of chunk c, so memory latency overlaps the matrix pipe.
// K steps (of 4 gathered rows) per pipeline stage: FP64 fragments are twice
// as wide in registers, so they stage half as many rows.
template <typename T>
struct Stage {
  static constexpr int KS = 4;
};
template <>
struct Stage<double> {
  static constexpr int KS = 2;
};

template <typename T>
struct Frag {
  T a[Stage<T>::KS][4];  // [k step][x block]
  T b[Stage<T>::KS][4];  // [k step][y block]
};

template <typename T>
__device__ inline void load_chunk(Frag<T>& f, const T* __restrict__ A, int64_t X, const uint32_t* __restrict__ rowlist,
                                  const T* __restrict__ B, uint32_t kc, const int64_t xa[4], int lr, int lk) {
#pragma unroll
  for (int j = 0; j < Stage<T>::KS; j++) {
    const uint32_t kr = kc + 4 * j + lk;
    const T* ar = A + (int64_t)rowlist[kr] * X;
#pragma unroll
    for (int s = 0; s < 4; s++) f.a[j][s] = ar[xa[s]];
    const T* br = B + (int64_t)kr * GT_Y + lr;
#pragma unroll
    for (int q = 0; q < 4; q++) f.b[j][q] = br[16 * q];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_ggemm(const T* __restrict__ A, int64_t X, const uint32_t* __restrict__ tile_off,
                                               const uint32_t* __restrict__ rowlist, const T* __restrict__ B,
                                               int64_t nbx, int64_t ntiles, int64_t Y, OutDev om, double alpha, int order) {
  typedef typename Acc<T>::type AT;
  int64_t bx, t;
  xcd_remap(nbx, ntiles, order, bx, t);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t xw = bx * GT_X + 64 * w;
  const uint32_t k_beg = tile_off[t], k_end = tile_off[t + 1];
  AT acc[4][4];
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int q = 0; q < 4; q++) acc[s][q] = AT{0, 0, 0, 0};
  int64_t xa[4];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    xa[s] = xw + 16 * s + lr;
    if (xa[s] >= X) xa[s] = X - 1;
  }
  constexpr uint32_t KCH = 4 * Stage<T>::KS;  // rows per stage; views pad tiles to 16 rows
  if (k_beg < k_end) {
    Frag<T> f0, f1;
    load_chunk(f0, A, X, rowlist, B, k_beg, xa, lr, lk);
    uint32_t kc = k_beg;
    while (true) {
      const bool more = kc + KCH < k_end;
      if (more) load_chunk(f1, A, X, rowlist, B, kc + KCH, xa, lr, lk);
#pragma unroll
      for (int j = 0; j < Stage<T>::KS; j++)
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
          for (int q = 0; q < 4; q++) acc[s][q] = Acc<T>::mfma(f0.a[j][s], f0.b[j][q], acc[s][q]);
      if (!more) break;
      kc += KCH;
      const bool more2 = kc + KCH < k_end;
      if (more2) load_chunk(f0, A, X, rowlist, B, kc + KCH, xa, lr, lk);
#pragma unroll
      for (int j = 0; j < Stage<T>::KS; j++)
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
          for (int q = 0; q < 4; q++) acc[s][q] = Acc<T>::mfma(f1.a[j][s], f1.b[j][q], acc[s][q]);
      if (!more2) break;
      kc += KCH;
    }
  }
  // D layout: lane holds column lane%16 of each 16x16 block, rows Acc<T>::row(lane/16, i)
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int64_t y = t * GT_Y + 16 * q + lr;
    if (y >= Y) continue;
    const uint32_t yy = (uint32_t)y;
    const int64_t oy = (int64_t)(yy / om.Yl) * om.oyh + (int64_t)(yy % om.Yl) * om.oyl;
#pragma unroll
    for (int s = 0; s < 4; s++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int64_t x = xw + 16 * s + Acc<T>::row(lk, i);
        if (x < X) {
          const uint32_t xx = (uint32_t)x;
          out_add(om, (int64_t)(xx / om.Xl) * om.oxh + (int64_t)(xx % om.Xl) * om.oxl + oy, alpha,
                  (double)acc[s][q][i]);
        }
      }
    }
  }
}

int g_order = 1;

int gathered_gemm(const View* v, const void* A, int adtype, int64_t X, const OutMap& m, double alpha,
                  hipStream_t s) {
  FRICC_CHECK(v->TY == GT_Y, "gathered_gemm: view tile width must be 64");
  FRICC_CHECK(v->dtype == adtype, "gathered_gemm: operand dtypes differ");
  FRICC_CHECK(X > 0 && X < ((int64_t)1 << 31), "gathered_gemm: bad X");
  if (v->ntiles == 0) return OK;
  OutDev om;
  om.out = m.out;
  om.of32 = m.of32;
  om.Xl = (uint32_t)m.Xl;
  om.Yl = (uint32_t)v->Yl;
  om.oxh = m.oxh;
  om.oxl = m.oxl;
  om.oyh = m.oyh;
  om.oyl = m.oyl;
  const int64_t nbx = ceil_div(X, GT_X);
  const int64_t nb = nbx * v->ntiles;
  if (adtype == DT_F64) {
    k_ggemm<double><<<nb, 256, 0, s>>>((const double*)A, X, v->tile_off, v->rowlist, (const double*)v->B, nbx,
                                       v->ntiles, v->Y, om, alpha, g_order);
  } else if (adtype == DT_F32) {
    k_ggemm<float><<<nb, 256, 0, s>>>((const float*)A, X, v->tile_off, v->rowlist, (const float*)v->B, nbx,
                                      v->ntiles, v->Y, om, alpha, g_order);
  } else {
    FRICC_CHECK(false, "gathered_gemm: unsupported dtype");
  }
  FRICC_HIP(hipGetLastError());
  return OK;
}

// ---------------------------------------------------------------------------
// BF16 gathered contraction (v_mfma_f32_16x16x32_bf16, FP32 accumulate).
// A and B come as hi/lo bf16 planes; the 8 rows of a group are contiguous
// (16 bytes) so each lane loads its 8 K values of one x (or y) at once.
// npass 3: hi*hi + hi*lo + lo*hi (about 16 significant bits per product).
typedef __attribute__((ext_vector_type(8))) __bf16 bf8;

__device__ inline bf8 as_bf8(uint4 u) {
  bf8 r;
  __builtin_memcpy(&r, &u, 16);
  return r;
}

template <int NPASS>
struct FragB16 {
  uint4 ah[4], al[4], bh[4], bl[4];
};

template <int NPASS>
__device__ inline void load_b16(FragB16<NPASS>& f, const uint4* __restrict__ Ahi, const uint4* __restrict__ Alo,
                                int64_t X, const uint32_t* __restrict__ glist, const uint4* __restrict__ Bhi,
                                const uint4* __restrict__ Blo, uint32_t kg, const int64_t xa[4], int lr) {
  const int64_t ga = (int64_t)glist[kg] * X;
  const int64_t gb = (int64_t)kg * GT_Y + lr;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    f.ah[s] = Ahi[ga + xa[s]];
    if (NPASS > 1) f.al[s] = Alo[ga + xa[s]];
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    f.bh[q] = Bhi[gb + 16 * q];
    if (NPASS > 1) f.bl[q] = Blo[gb + 16 * q];
  }
}

template <int NPASS>
__device__ inline void mma_b16(f4 (&acc)[4][4], const FragB16<NPASS>& f) {
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      acc[s][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(f.ah[s]), as_bf8(f.bh[q]), acc[s][q], 0, 0, 0);
      if (NPASS > 1) {
        acc[s][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(f.ah[s]), as_bf8(f.bl[q]), acc[s][q], 0, 0, 0);
        acc[s][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(f.al[s]), as_bf8(f.bh[q]), acc[s][q], 0, 0, 0);
      }
    }
}

template <int NPASS>
__global__ __launch_bounds__(256) void k_ggemm_bf16(const uint4* __restrict__ Ahi, const uint4* __restrict__ Alo,
                                                    int64_t X, const uint32_t* __restrict__ tile_off,
                                                    const uint32_t* __restrict__ glist, const uint4* __restrict__ Bhi,
                                                    const uint4* __restrict__ Blo, int64_t nbx, int64_t ntiles,
                                                    int64_t Y, OutDev om, double alpha, int order) {
  int64_t bx, t;
  xcd_remap(nbx, ntiles, order, bx, t);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t xw = bx * GT_X + 64 * w;
  const uint32_t g_beg = tile_off[t], g_end = tile_off[t + 1];  // groups; multiple of 4 per tile
  f4 acc[4][4];
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int q = 0; q < 4; q++) acc[s][q] = f4{0, 0, 0, 0};
  int64_t xa[4];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    xa[s] = xw + 16 * s + lr;
    if (xa[s] >= X) xa[s] = X - 1;
  }
  if (g_beg < g_end) {
    FragB16<NPASS> f0, f1;
    load_b16<NPASS>(f0, Ahi, Alo, X, glist, Bhi, Blo, g_beg + lk, xa, lr);
    uint32_t kg = g_beg;
    while (true) {
      const bool more = kg + 4 < g_end;
      if (more) load_b16<NPASS>(f1, Ahi, Alo, X, glist, Bhi, Blo, kg + 4 + lk, xa, lr);
      mma_b16<NPASS>(acc, f0);
      if (!more) break;
      kg += 4;
      const bool more2 = kg + 4 < g_end;
      if (more2) load_b16<NPASS>(f0, Ahi, Alo, X, glist, Bhi, Blo, kg + 4 + lk, xa, lr);
      mma_b16<NPASS>(acc, f1);
      if (!more2) break;
      kg += 4;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int64_t y = t * GT_Y + 16 * q + lr;
    if (y >= Y) continue;
    const uint32_t yy = (uint32_t)y;
    const int64_t oy = (int64_t)(yy / om.Yl) * om.oyh + (int64_t)(yy % om.Yl) * om.oyl;
#pragma unroll
    for (int s = 0; s < 4; s++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int64_t x = xw + 16 * s + 4 * lk + i;
        if (x < X) {
          const uint32_t xx = (uint32_t)x;
          out_add(om, (int64_t)(xx / om.Xl) * om.oxh + (int64_t)(xx % om.Xl) * om.oxl + oy, alpha,
                  (double)acc[s][q][i]);
        }
      }
    }
  }
}

int gathered_gemm_bf16(const View* v, const uint16_t* A2, int64_t R, int64_t X, const OutMap& m, double alpha,
                       int npass, hipStream_t s) {
  FRICC_CHECK(v->TY == GT_Y && v->dtype == DT_BF16, "gathered_gemm_bf16: needs a BF16 view with tile width 64");
  FRICC_CHECK(X > 0 && X < ((int64_t)1 << 31), "gathered_gemm_bf16: bad X");
  FRICC_CHECK(npass == 1 || npass == 3, "gathered_gemm_bf16: npass must be 1 or 3");
  if (v->ntiles == 0) return OK;
  OutDev om;
  om.out = m.out;
  om.of32 = m.of32;
  om.Xl = (uint32_t)m.Xl;
  om.Yl = (uint32_t)v->Yl;
  om.oxh = m.oxh;
  om.oxl = m.oxl;
  om.oyh = m.oyh;
  om.oyl = m.oyl;
  const int64_t Rg = (R + 7) / 8;
  const uint4* Ahi = (const uint4*)A2;
  const uint4* Alo = Ahi + Rg * X;
  const int64_t nbx = ceil_div(X, GT_X);
  const int64_t nb = nbx * v->ntiles;
  if (npass == 3)
    k_ggemm_bf16<3><<<nb, 256, 0, s>>>(Ahi, Alo, X, v->tile_off, v->rowlist, (const uint4*)v->B, (const uint4*)v->Blo,
                                       nbx, v->ntiles, v->Y, om, alpha, g_order);
  else
    k_ggemm_bf16<1><<<nb, 256, 0, s>>>(Ahi, Alo, X, v->tile_off, v->rowlist, (const uint4*)v->B, (const uint4*)v->Blo,
                                       nbx, v->ntiles, v->Y, om, alpha, g_order);
  FRICC_HIP(hipGetLastError());
  return OK;
}

// A[R][X] (f64 or f32) -> hi/lo planes [(R+7)/8][X][8]; rows past R are zero.
template <typename T>
__global__ void k_split_bf16(const T* __restrict__ A, int64_t R, int64_t X, __hip_bfloat16* __restrict__ hi,
                             __hip_bfloat16* __restrict__ lo) {
  const int64_t gx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t Rg = (R + 7) / 8;
  if (gx >= Rg * X) return;
  const int64_t g = gx / X, x = gx % X;
  __hip_bfloat16 h[8], l[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int64_t r = g * 8 + j;
    const float v = r < R ? (float)A[r * X + x] : 0.0f;
    h[j] = __float2bfloat16(v);
    l[j] = __float2bfloat16(v - __bfloat162float(h[j]));
  }
  uint4 uh, ul;
  __builtin_memcpy(&uh, h, 16);
  __builtin_memcpy(&ul, l, 16);
  ((uint4*)hi)[gx] = uh;
  ((uint4*)lo)[gx] = ul;
}

int split_bf16(const void* A, int src_dtype, int64_t R, int64_t X, uint16_t* A2, hipStream_t s) {
  const int64_t Rg = (R + 7) / 8;
  const int64_t n = Rg * X;
  if (n == 0) return OK;
  __hip_bfloat16* hi = (__hip_bfloat16*)A2;
  __hip_bfloat16* lo = hi + n * 8;
  if (src_dtype == DT_F64)
    k_split_bf16<double><<<ceil_div(n, 256), 256, 0, s>>>((const double*)A, R, X, hi, lo);
  else if (src_dtype == DT_F32)
    k_split_bf16<float><<<ceil_div(n, 256), 256, 0, s>>>((const float*)A, R, X, hi, lo);
  else
    FRICC_CHECK(false, "split_bf16: source must be FP64 or FP32");
  FRICC_HIP(hipGetLastError());
  return OK;
}

// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_permute(const double* __restrict__ W, int64_t Rl, int64_t Xh, int64_t Xl, int64_t s0, int64_t s1,
                          int64_t s2, int64_t s3, T* __restrict__ A) {
  const int64_t r = blockIdx.x;
  const int64_t rh = r / Rl, rl = r % Rl;
  const int64_t X = Xh * Xl;
  const double* wr = W + rh * s0 + rl * s1;
  T* ar = A + r * X;
  for (int64_t x = threadIdx.x; x < X; x += blockDim.x) {
    const int64_t xh = x / Xl, xl = x % Xl;
    ar[x] = (T)wr[xh * s2 + xl * s3];
  }
}

int permute_operand(const double* W, int64_t Rh, int64_t Rl, int64_t Xh, int64_t Xl, const int64_t st[4], void* A,
                    int dtype, hipStream_t s) {
  const int64_t R = Rh * Rl;
  if (R == 0 || Xh * Xl == 0) return OK;
  if (dtype == DT_F64)
    k_permute<double><<<R, 256, 0, s>>>(W, Rl, Xh, Xl, st[0], st[1], st[2], st[3], (double*)A);
  else if (dtype == DT_F32)
    k_permute<float><<<R, 256, 0, s>>>(W, Rl, Xh, Xl, st[0], st[1], st[2], st[3], (float*)A);
  else
    FRICC_CHECK(false, "permute_operand: unsupported dtype");
  FRICC_HIP(hipGetLastError());
  return OK;
}

}  // namespace fricc
