// Shared device/host helpers for libfricc_hip (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <string.h>

#include <string>

namespace fricc {

typedef unsigned __int128 u128;

// ---------------------------------------------------------------------------
// Error plumbing: every entry point returns an int status and records a
// message retrievable through fricc_last_error().
void set_error(const std::string& msg);

enum Status {
  OK = 0,
  EINVAL_ = 1,
  EDEVICE = 2,
  ENOMEM_ = 3,
  EPROB = 4,      // some p_i > 1 (the reference exits, src/fri_utils.cpp:383-395)
  ENONFINITE = 5, // non-finite entry in the vector to compress
  ESAMPLE = 6,    // pivotal sampling ends with the wrong sample count (the reference exits)
};

#define FRICC_HIP(call)                                                          \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      fricc::set_error(std::string(#call) + ": " + hipGetErrorString(e_));      \
      return fricc::EDEVICE;                                                     \
    }                                                                            \
  } while (0)

#define FRICC_CHECK(cond, msg)                                                   \
  do {                                                                           \
    if (!(cond)) {                                                               \
      fricc::set_error(msg);                                                     \
      return fricc::EINVAL_;                                                     \
    }                                                                            \
  } while (0)

#define FRICC_TRY(expr)                                                          \
  do {                                                                           \
    int s_ = (expr);                                                             \
    if (s_ != 0) return s_;                                                      \
  } while (0)

// ---------------------------------------------------------------------------
// Exact fixed-point arithmetic used for every sum that feeds a sampling
// decision.  The grid for |x| is 2^G, G = max(eb_max,1) - 1119, where eb_max
// is the biased exponent of max|x|: values within 2^44 of the top binade are
// represented exactly, smaller ones are truncated (relative 2^-96).  The grid
// for probabilities p in [0,1] is fixed at 2^-96.  Integer sums are
// associative, so every parallel order gives the same bits.

__host__ __device__ inline uint32_t exp_field(uint64_t bits) { return (uint32_t)((bits >> 52) & 0x7ff); }

__host__ __device__ inline u128 fx_of(uint64_t ybits, uint32_t eb_max) {
  uint32_t eb = exp_field(ybits);
  uint64_t M = (ybits & ((1ull << 52) - 1)) | (eb ? (1ull << 52) : 0ull);
  int s = (int)(eb > 1u ? eb : 1u) - (int)(eb_max > 1u ? eb_max : 1u) + 44;
  if (eb == 0x7ffu) return 0;  // non-finite: callers reject these separately
  if (s > 74) s = 74;          // only reachable for eb > eb_max (invalid use); keeps < 2^127
  if (s >= 0) return ((u128)M) << s;
  int r = -s;
  return r >= 64 ? (u128)0 : (u128)(M >> r);
}

__host__ __device__ inline int fx_grid(uint32_t eb_max) { return (int)(eb_max > 1u ? eb_max : 1u) - 1119; }

__host__ __device__ inline int clz64(uint64_t v) {
#ifdef __HIP_DEVICE_COMPILE__
  return __clzll(v);
#else
  return __builtin_clzll(v);
#endif
}

// Round-to-nearest-even value of v * 2^G.
__host__ __device__ inline double u128_to_double(u128 v, int G) {
  if (v == 0) return 0.0;
  uint64_t hi = (uint64_t)(v >> 64), lo = (uint64_t)v;
  int nb = hi ? 128 - clz64(hi) : 64 - clz64(lo);
  uint64_t mant;
  int e = 0;
  if (nb <= 53) {
    mant = lo;
  } else {
    int sh = nb - 53;
    mant = (uint64_t)(v >> sh);
    u128 rem = v & ((((u128)1) << sh) - 1);
    u128 half = ((u128)1) << (sh - 1);
    if (rem > half || (rem == half && (mant & 1ull))) {
      mant++;
      if (mant == (1ull << 53)) {
        mant >>= 1;
        sh++;
      }
    }
    e = sh;
  }
  return ldexp((double)mant, e + G);
}

__host__ __device__ inline uint64_t dbits(double v) {
#ifdef __HIP_DEVICE_COMPILE__
  return (uint64_t)__double_as_longlong(v);
#else
  uint64_t b;
  memcpy(&b, &v, 8);
  return b;
#endif
}

// Four 32-bit limbs accumulated in 64-bit words: exact, carry-free for up to
// 2^31 addends of < 2^97, and combinable with plain 64-bit atomics.
struct Limbs {
  uint64_t l[4];
};

__device__ inline void limbs_add(uint64_t* L, u128 v) {
  L[0] += (uint64_t)(uint32_t)v;
  L[1] += (uint64_t)(uint32_t)(v >> 32);
  L[2] += (uint64_t)(uint32_t)(v >> 64);
  L[3] += (uint64_t)(uint32_t)(v >> 96);
}

__host__ __device__ inline u128 limbs_value(const uint64_t* L) {
  return (u128)L[0] + ((u128)L[1] << 32) + ((u128)L[2] << 64) + ((u128)L[3] << 96);
}

// ---------------------------------------------------------------------------
// Wavefront helpers (wave64).
constexpr int WAVE = 64;

__device__ inline uint64_t shfl_up_u64(uint64_t v, int d) {
  return (uint64_t)__shfl_up((unsigned long long)v, d, WAVE);
}
__device__ inline uint64_t shfl_u64(uint64_t v, int src) {
  return (uint64_t)__shfl((unsigned long long)v, src, WAVE);
}
__device__ inline u128 shfl_up_u128(u128 v, int d) {
  uint64_t lo = shfl_up_u64((uint64_t)v, d), hi = shfl_up_u64((uint64_t)(v >> 64), d);
  return ((u128)hi << 64) | lo;
}
__device__ inline u128 shfl_u128(u128 v, int src) {
  uint64_t lo = shfl_u64((uint64_t)v, src), hi = shfl_u64((uint64_t)(v >> 64), src);
  return ((u128)hi << 64) | lo;
}

// Inclusive wave scan of u128 (exact).
__device__ inline u128 wave_incl_scan_u128(u128 v) {
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    u128 t = shfl_up_u128(v, d);
    if (lane >= d) v += t;
  }
  return v;
}

__device__ inline uint32_t wave_incl_scan_u32(uint32_t v) {
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    uint32_t t = __shfl_up(v, d, WAVE);
    if (lane >= d) v += t;
  }
  return v;
}

// Block-wide exclusive scan of u128 for blockDim.x == NT (multiple of 64).
// Returns the exclusive prefix of this thread and writes the block total.
template <int NT>
__device__ inline u128 block_excl_scan_u128(u128 v, u128* total, u128* smem /*NT/64+1*/) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  u128 inc = wave_incl_scan_u128(v);
  if (lane == WAVE - 1) smem[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    u128 run = 0;
    for (int i = 0; i < NT / WAVE; i++) {
      u128 t = smem[i];
      smem[i] = run;
      run += t;
    }
    smem[NT / WAVE] = run;
  }
  __syncthreads();
  u128 ex = smem[w] + inc - v;
  *total = smem[NT / WAVE];
  __syncthreads();
  return ex;
}

template <int NT>
__device__ inline uint32_t block_excl_scan_u32(uint32_t v, uint32_t* total, uint32_t* smem /*NT/64+1*/) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  uint32_t inc = wave_incl_scan_u32(v);
  if (lane == WAVE - 1) smem[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int i = 0; i < NT / WAVE; i++) {
      uint32_t t = smem[i];
      smem[i] = run;
      run += t;
    }
    smem[NT / WAVE] = run;
  }
  __syncthreads();
  uint32_t ex = smem[w] + inc - v;
  *total = smem[NT / WAVE];
  __syncthreads();
  return ex;
}

// Block-wide exclusive max-scan of u32 (0 before the first thread) for
// blockDim.x == NT; writes the block maximum.
template <int NT>
__device__ inline uint32_t block_excl_max_u32(uint32_t v, uint32_t* total, uint32_t* smem /*NT/64+1*/) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    uint32_t t = __shfl_up(inc, d, WAVE);
    if (lane >= d) inc = inc > t ? inc : t;
  }
  uint32_t ex = __shfl_up(inc, 1, WAVE);
  if (lane == 0) ex = 0;
  if (lane == WAVE - 1) smem[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int i = 0; i < NT / WAVE; i++) {
      uint32_t t = smem[i];
      smem[i] = run;
      run = run > t ? run : t;
    }
    smem[NT / WAVE] = run;
  }
  __syncthreads();
  ex = ex > smem[w] ? ex : smem[w];
  *total = smem[NT / WAVE];
  __syncthreads();
  return ex;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// Device workspace: grow-only slots of the calling thread's selected arena on
// the current device (arena.h holds the rules).
int ws_get(int slot, size_t bytes, void** out);

// Exclusive scan of n uint32 values (in place allowed), via tile sums.  sums:
// scan_u32_sums_words(n) words of device scratch, or nullptr to take them from
// the workspace.
int64_t scan_u32_sums_words(int64_t n);
int scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev, hipStream_t s,
             uint32_t* sums = nullptr);

}  // namespace fricc
