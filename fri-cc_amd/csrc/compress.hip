// FRI compression of a flat FP64 vector on the GPU (gfx950).
//
// Replaces fri_compression / get_d_largest / one_norm / argsort /
// make_probability_vector / sample_systematic (src/fri_utils.cpp:11-36,
// 45-58, 129-171, 367-522) and the "largest" path of the SparseTensor4d
// constructor (include/sparse_tensor.hpp:46-56).
//
// Canonical semantics (identical, bit for bit, to the CPU oracle's canonical
// mode):
//   sigma   = indices by |x| descending, ties by ascending index
//   R0      = round(sum FX(|x_i|))                     exact fixed point
//   D       = longest prefix of sigma (k < m) with
//             (m-k)*y_k >= R_k - y_k && R_k > 1e-14,    R_k = round(Ftot - sum_{l<k} FX(y_l))
//   p_j     = (j in D) ? 0 : |x_j| / R_|D| * n'         n' = m - |D|
//   P_j     = round(2^-96 * sum_{l<j} floor(p_l * 2^96))
//   S_k     = j with P_j <= fl(k + u) < P_{j+1}, or 0 if no such j
//             (systematic), or the pivotal chain's k-th emission (pivotal,
//             src/fri_utils.cpp:63-127, on the exact prefix of FXP(p))
//   output  = [sigma_0..sigma_{|D|-1}, S_0..S_{n'-1}],
//             values [x_D, x_S / p_S]
//
// Pipeline (one stream, no host round trip; fricc_compress reads the scalars
// back once at the end):
//   A  histogram of the top 12 bits of |x| (+ non-finite detection, max exponent)
//   sel choose the top-12-bit bucket b1 holding the m-th largest
//   B  exact FX 1-norm + per-tile count of candidates (bucket >= b1)
//   C  order-preserving compaction of the candidates (tile scan + scatter)
//   sort stable LSD radix sort of the candidates by |x| descending
//   D  exact prefix of FX over the sorted candidates, parallel test, |D| by atomicMin
//   S  exact prefix of FXP(p) over all N, element-parallel systematic emission
#include "common.h"
#include "fricc_internal.h"

#include <mutex>

namespace fricc {

constexpr int NT = 256;
constexpr int ITEMS = 16;
constexpr int TILE = NT * ITEMS;  // elements per tile, striped rounds of NT
constexpr uint64_t MAXKEY = 0x7fffffffffffffffull;

struct CompScalars {
  uint32_t ebmax, nonfinite, b1, pflag;
  uint64_t n_cand;
  uint64_t limbs[4];
  unsigned long long K;
  uint64_t n_sampled;
  double R0, RK, Ptot;
  uint64_t thr_key, thr_idx;
  uint64_t k_notfound;
  uint64_t n_notfound;
  double u;        // the uniform of this call
  uint32_t ubad;   // u outside [0, 1]
  uint32_t dup;    // the output may repeat an index (two samples in one interval, or samples past the end)
  uint64_t key;    // pivotal: SplitMix64 key of the step draws (the bits of u)
  uint32_t pivbad; // pivotal: the chain ends with a sample count other than n'
  uint32_t pad2;
};

// ---------------------------------------------------------------- pass A
__global__ void k_hist_top(const double* __restrict__ x, int64_t n, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[4096];
  for (int i = threadIdx.x; i < 4096; i += NT) h[i] = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) {
    uint64_t yb = dbits(x[i]) & MAXKEY;
    atomicAdd(&h[yb >> 51], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 4096; i += NT)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// Single block: bucket b1 with count(>b1) < m <= count(>=b1).
__global__ void k_select_top(const uint32_t* __restrict__ hist, int64_t m, CompScalars* sc) {
  __shared__ uint32_t sm[NT / WAVE + 1];
  __shared__ int s_top;
  if (threadIdx.x == 0) s_top = -1;
  __syncthreads();
  // thread t owns buckets [4095 - 16t - 15, 4095 - 16t] (descending)
  uint32_t c[16];
  uint32_t loc = 0;
  for (int q = 0; q < 16; q++) {
    int b = 4095 - (threadIdx.x * 16 + q);
    c[q] = hist[b];
    loc += c[q];
  }
  for (int q = 0; q < 16; q++)
    if (c[q]) {
      atomicMax(&s_top, 4095 - (int)(threadIdx.x * 16 + q));
      break;
    }
  uint32_t tot;
  uint32_t above = block_excl_scan_u32<NT>(loc, &tot, sm);  // count strictly above my range
  if ((int64_t)above < m && (int64_t)(above + loc) >= m) {
    uint64_t run = above;
    for (int q = 0; q < 16; q++) {
      if ((int64_t)(run + c[q]) >= m) {
        sc->b1 = 4095 - (threadIdx.x * 16 + q);
        sc->n_cand = run + c[q];
        break;
      }
      run += c[q];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int top = s_top < 0 ? 0 : s_top;
    sc->ebmax = (uint32_t)top >> 1;
    sc->nonfinite = (top >= 0xffe) ? 1u : 0u;
  }
}

// ---------------------------------------------------------------- pass B
__global__ void k_norm_count(const double* __restrict__ x, int64_t n, CompScalars* sc,
                             uint32_t* __restrict__ tilecnt) {
  __shared__ uint64_t sl[4][NT / WAVE];
  __shared__ uint32_t sm[NT / WAVE + 1];
  const uint32_t ebmax = sc->ebmax, b1 = sc->b1;
  const int64_t base = (int64_t)blockIdx.x * TILE;
  uint64_t L[4] = {0, 0, 0, 0};
  uint32_t cnt = 0;
#pragma unroll 4
  for (int r = 0; r < ITEMS; r++) {
    int64_t i = base + (int64_t)r * NT + threadIdx.x;
    if (i < n) {
      uint64_t yb = dbits(x[i]) & MAXKEY;
      limbs_add(L, fx_of(yb, ebmax));
      cnt += (uint32_t)((yb >> 51) >= b1);
    }
  }
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  for (int q = 0; q < 4; q++) {
    uint64_t v = L[q];
    for (int d = WAVE / 2; d > 0; d >>= 1) v += (uint64_t)__shfl_down((unsigned long long)v, d, WAVE);
    if (lane == 0) sl[q][w] = v;
  }
  uint32_t tot;
  block_excl_scan_u32<NT>(cnt, &tot, sm);
  if (threadIdx.x < 4) {
    uint64_t v = 0;
    for (int k = 0; k < NT / WAVE; k++) v += sl[threadIdx.x][k];
    atomicAdd((unsigned long long*)&sc->limbs[threadIdx.x], (unsigned long long)v);
  }
  if (threadIdx.x == 0) tilecnt[blockIdx.x] = tot;
}

// ---------------------------------------------------------------- pass C
// Order-preserving compaction of candidates: key = MAXKEY - |x| bits (ascending
// key == descending |x|), value = index.
__global__ void k_compact(const double* __restrict__ x, int64_t n, const CompScalars* __restrict__ sc,
                          const uint32_t* __restrict__ tileoff, uint64_t* __restrict__ keys,
                          uint32_t* __restrict__ vals) {
  __shared__ uint32_t wc[NT / WAVE];
  __shared__ uint32_t s_run;
  const uint32_t b1 = sc->b1;
  const int64_t base = (int64_t)blockIdx.x * TILE;
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  if (threadIdx.x == 0) s_run = tileoff[blockIdx.x];
  __syncthreads();
  for (int r = 0; r < ITEMS; r++) {
    int64_t i = base + (int64_t)r * NT + threadIdx.x;
    uint64_t yb = 0;
    bool pred = false;
    if (i < n) {
      yb = dbits(x[i]) & MAXKEY;
      pred = (yb >> 51) >= b1;
    }
    uint64_t bal = __ballot(pred);
    uint32_t rank = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wc[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t off = s_run;
    for (int k = 0; k < w; k++) off += wc[k];
    if (pred) {
      keys[off + rank] = MAXKEY - yb;
      vals[off + rank] = (uint32_t)i;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (int k = 0; k < NT / WAVE; k++) t += wc[k];
      s_run += t;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- radix sort
// Stable LSD radix sort of the n_cand candidates (count read on the device:
// no host round trip).  A fixed grid strides over the live tiles; per digit
// the per-tile counts live in blkhist[digit][tile] (row length ntmax).
__device__ inline int64_t live_tiles(const CompScalars* sc) { return ((int64_t)sc->n_cand + TILE - 1) / TILE; }

__global__ void k_sort_count(const uint64_t* __restrict__ keys, const CompScalars* __restrict__ sc, int shift,
                             uint32_t* __restrict__ blkhist, int64_t ntmax) {
  __shared__ uint32_t h[256];
  const int64_t n = (int64_t)sc->n_cand, nta = live_tiles(sc);
  for (int64_t t = blockIdx.x; t < nta; t += gridDim.x) {
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = t * TILE;
    for (int r = 0; r < ITEMS; r++) {
      int64_t i = base + (int64_t)r * NT + threadIdx.x;
      if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1u);
    }
    __syncthreads();
    blkhist[(int64_t)threadIdx.x * ntmax + t] = h[threadIdx.x];
    __syncthreads();
  }
}

// Block d: total of digit d over the live tiles.
__global__ void k_digit_total(const uint32_t* __restrict__ blkhist, int64_t ntmax, const CompScalars* __restrict__ sc,
                              uint32_t* __restrict__ dtot) {
  __shared__ uint32_t sm[NT / WAVE + 1];
  const int64_t nta = live_tiles(sc);
  const uint32_t* row = blkhist + (int64_t)blockIdx.x * ntmax;
  uint32_t s = 0;
  for (int64_t t = threadIdx.x; t < nta; t += NT) s += row[t];
  uint32_t tot;
  block_excl_scan_u32<NT>(s, &tot, sm);
  if (threadIdx.x == 0) dtot[blockIdx.x] = tot;
}

// Block d: exclusive offsets of digit d per live tile, after all smaller digits.
__global__ void k_digit_offsets(uint32_t* __restrict__ blkhist, int64_t ntmax, const CompScalars* __restrict__ sc,
                                const uint32_t* __restrict__ dtot) {
  __shared__ uint32_t sm[NT / WAVE + 1];
  const int64_t nta = live_tiles(sc);
  // start of digit blockIdx.x: the totals of all smaller digits
  uint32_t run;
  block_excl_scan_u32<NT>(threadIdx.x < blockIdx.x ? dtot[threadIdx.x] : 0u, &run, sm);
  uint32_t* row = blkhist + (int64_t)blockIdx.x * ntmax;
  for (int64_t c0 = 0; c0 < nta; c0 += NT) {
    const int64_t t = c0 + threadIdx.x;
    const uint32_t v = t < nta ? row[t] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan_u32<NT>(v, &tot, sm);
    if (t < nta) row[t] = run + ex;
    run += tot;
    __syncthreads();
  }
}

// Stable scatter: tile order = index order (rounds of NT, wave order, lane
// order).  Each tile is first put in digit order in LDS (stable), then written
// out so that consecutive threads write consecutive places of a digit's run.
__global__ __launch_bounds__(NT) void k_sort_scatter(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                     uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                     const CompScalars* __restrict__ sc, int shift,
                                                     const uint32_t* __restrict__ blkoff, int64_t ntmax) {
  __shared__ uint64_t sk[TILE];
  __shared__ uint32_t sv[TILE];
  __shared__ uint32_t gbase[256];   // this tile's first place in each digit's run (global)
  __shared__ uint32_t lstart[256];  // first place of each digit in the tile's digit order
  __shared__ uint32_t lrun[256];    // places of each digit taken so far
  __shared__ uint32_t wcnt[NT / WAVE][256];
  __shared__ uint32_t sm[NT / WAVE + 1];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int64_t n = (int64_t)sc->n_cand, nta = live_tiles(sc);
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int64_t t = blockIdx.x; t < nta; t += gridDim.x) {
    const int64_t base = t * TILE;
    const int cnt = (int)(n - base < TILE ? n - base : TILE);
    gbase[threadIdx.x] = blkoff[(int64_t)threadIdx.x * ntmax + t];
    lrun[threadIdx.x] = 0;
    uint64_t k[ITEMS];
    uint32_t v[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
      const int e = r * NT + threadIdx.x;
      k[r] = e < cnt ? kin[base + e] : 0;
      v[r] = e < cnt ? vin[base + e] : 0;
    }
    // digit counts of the tile -> local starts
    wcnt[0][threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; r++)
      if (r * NT + (int)threadIdx.x < cnt) atomicAdd(&wcnt[0][(k[r] >> shift) & 255], 1u);
    __syncthreads();
    uint32_t tot;
    lstart[threadIdx.x] = block_excl_scan_u32<NT>(wcnt[0][threadIdx.x], &tot, sm);
    // stable placement, one round of NT elements at a time
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
      for (int q = threadIdx.x; q < (NT / WAVE) * 256; q += NT) (&wcnt[0][0])[q] = 0;
      __syncthreads();
      const bool valid = r * NT + (int)threadIdx.x < cnt;
      const uint32_t d = (uint32_t)((k[r] >> shift) & 255);
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const bool bit = (d >> b) & 1;
        const uint64_t bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
      }
      const uint32_t rank = (uint32_t)__popcll(peers & lt);
      if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
      __syncthreads();
      if (valid) {
        uint32_t pos = lstart[d] + lrun[d] + rank;
        for (int q = 0; q < w; q++) pos += wcnt[q][d];
        sk[pos] = k[r];
        sv[pos] = v[r];
      }
      __syncthreads();
      uint32_t add = 0;
      for (int q = 0; q < NT / WAVE; q++) add += wcnt[q][threadIdx.x];
      lrun[threadIdx.x] += add;
      __syncthreads();
    }
    // write out in digit order
    for (int p = threadIdx.x; p < cnt; p += NT) {
      const uint64_t kk = sk[p];
      const uint32_t d = (uint32_t)((kk >> shift) & 255);
      const uint32_t o = gbase[d] + (uint32_t)p - lstart[d];
      kout[o] = kk;
      vout[o] = sv[p];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- D test
__device__ inline double key_to_y(uint64_t key) {
  return __longlong_as_double((long long)(MAXKEY - key));
}

__global__ void k_dfx_tilesums(const uint64_t* __restrict__ keys, int64_t m, const CompScalars* __restrict__ sc,
                               u128* __restrict__ tsum) {
  __shared__ u128 sm[NT / WAVE + 1];
  const uint32_t ebmax = sc->ebmax;
  const int64_t base = (int64_t)blockIdx.x * TILE + (int64_t)threadIdx.x * ITEMS;
  u128 s = 0;
  for (int r = 0; r < ITEMS; r++) {
    int64_t k = base + r;
    if (k < m) s += fx_of(MAXKEY - keys[k], ebmax);
  }
  u128 tot;
  block_excl_scan_u128<NT>(s, &tot, sm);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

// Single block, exclusive scan of u128 tile sums in place; sums[nt] = total.
__global__ void k_scan_u128_sums(u128* sums, int64_t nt, u128* total) {
  __shared__ u128 sm[1024 / WAVE + 1];
  const int64_t chunk = (nt + 1023) / 1024;
  const int64_t lo = threadIdx.x * chunk, hi = lo + chunk < nt ? lo + chunk : nt;
  u128 s = 0;
  for (int64_t i = lo; i < hi; i++) s += sums[i];
  u128 tot;
  u128 ex = block_excl_scan_u128<1024>(s, &tot, sm);
  for (int64_t i = lo; i < hi; i++) {
    u128 t = sums[i];
    sums[i] = ex;
    ex += t;
  }
  if (threadIdx.x == 0) {
    sums[nt] = tot;  // (nt + 1 entries) the end of the last tile
    if (total) *total = tot;
  }
}

__global__ void k_dtest(const uint64_t* __restrict__ keys, int64_t m, CompScalars* sc,
                        const u128* __restrict__ tsum) {
  __shared__ u128 sm[NT / WAVE + 1];
  const uint32_t ebmax = sc->ebmax;
  const int G = fx_grid(ebmax);
  const u128 Ftot = limbs_value(sc->limbs);
  const int64_t base = (int64_t)blockIdx.x * TILE + (int64_t)threadIdx.x * ITEMS;
  u128 f[ITEMS];
  u128 s = 0;
#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    int64_t k = base + r;
    f[r] = k < m ? fx_of(MAXKEY - keys[k], ebmax) : (u128)0;
    s += f[r];
  }
  u128 tot;
  u128 pre = block_excl_scan_u128<NT>(s, &tot, sm) + tsum[blockIdx.x];
  unsigned long long first_fail = (unsigned long long)m;
#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    int64_t k = base + r;
    if (k < m) {
      double Rk = u128_to_double(Ftot - pre, G);
      double y = key_to_y(keys[k]);
      bool pass = ((double)(m - k) * y >= Rk - y) && (Rk > 1e-14);
      if (!pass && (unsigned long long)k < first_fail) first_fail = (unsigned long long)k;
    }
    pre += f[r];
  }
  if (first_fail < (unsigned long long)m) atomicMin(&sc->K, first_fail);
}

// Single block: R_K, n', D threshold.
__global__ void k_dfinal(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t m,
                         CompScalars* sc, const u128* __restrict__ tsum) {
  __shared__ u128 sm[NT / WAVE + 1];
  const uint32_t ebmax = sc->ebmax;
  const int G = fx_grid(ebmax);
  const u128 Ftot = limbs_value(sc->limbs);
  const int64_t K = (int64_t)sc->K;
  // prefix of FX over sigma_0..sigma_{K-1}: tile prefix + partial sum inside
  // the tile holding position K (the last tile when K == m)
  const int64_t ntm = (m + TILE - 1) / TILE;
  const int64_t t = (K / TILE < ntm) ? K / TILE : ntm - 1;
  const int64_t base = t * TILE + (int64_t)threadIdx.x * ITEMS;
  u128 s = 0;
  for (int r = 0; r < ITEMS; r++) {
    int64_t k = base + r;
    if (k < K && k < m) s += fx_of(MAXKEY - keys[k], ebmax);
  }
  u128 tot;
  block_excl_scan_u128<NT>(s, &tot, sm);
  if (threadIdx.x == 0) {
    u128 pre = tsum[t] + tot;
    sc->R0 = u128_to_double(Ftot, G);
    sc->RK = u128_to_double(Ftot - pre, G);
    sc->n_sampled = (uint64_t)(m - K);
    if (K > 0) {
      sc->thr_key = keys[K - 1];
      sc->thr_idx = vals[K - 1];
    } else {
      sc->thr_key = 0;
      sc->thr_idx = 0;
    }
  }
}

// ---------------------------------------------------------------- sampling
__device__ inline double prob_of(double xv, uint64_t j, const CompScalars& c, double nsd) {
  uint64_t yb = dbits(xv) & MAXKEY;
  uint64_t key = MAXKEY - yb;
  bool inD = c.K > 0 && (key < c.thr_key || (key == c.thr_key && j <= c.thr_idx));
  return inD ? 0.0 : (fabs(xv) / c.RK) * nsd;
}

// One tile (TILE doubles) of src staged in LDS with coalesced loads; thread t
// then owns tile elements [t*ITEMS, t*ITEMS + ITEMS) (the blocked order the
// prefix needs).  Element e sits at LDS slot e + e/ITEMS: the pad slot keeps
// the blocked reads of a wave off one bank.
constexpr int LTILE = TILE + TILE / ITEMS;
__device__ inline void stage_tile(const double* __restrict__ src, int64_t base, int64_t n, double* lds) {
#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    const int e = r * NT + threadIdx.x;
    const int64_t j = base + e;
    lds[e + e / ITEMS] = j < n ? src[j] : 0.0;
  }
  __syncthreads();
}
__device__ inline double tile_elem(const double* lds, int r) {
  const int e = threadIdx.x * ITEMS + r;
  return lds[e + e / ITEMS];
}

// pin != nullptr: sample from the given probabilities (sample_systematic,
// src/fri_utils.cpp:129-159) instead of deriving them from x and D.
// Per tile: the exact sum of FXP(p) (tsum), and per thread the exact sum of
// FXP(p) over the tile's earlier threads (thpre), so that later passes find
// any thread's prefix without a block-wide scan.
__global__ void k_samp_tilesums(const double* __restrict__ x, const double* __restrict__ pin, int64_t n,
                                CompScalars* sc, u128* __restrict__ tsum, u128* __restrict__ thpre) {
  __shared__ u128 sm[NT / WAVE + 1];
  __shared__ double lds[LTILE];
  const CompScalars c = *sc;
  if (c.n_sampled == 0 && pin == nullptr) return;
  const double nsd = (double)c.n_sampled;
  const int64_t tb = (int64_t)blockIdx.x * TILE;
  const int64_t base = tb + (int64_t)threadIdx.x * ITEMS;
  stage_tile(pin ? pin : x, tb, n, lds);
  u128 s = 0;
  uint32_t bad = 0;
#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    const int64_t j = base + r;
    if (j < n) {
      const double v = tile_elem(lds, r);
      const double p = pin ? v : prob_of(v, (uint64_t)j, c, nsd);
      bad |= !(p >= 0.0 && p <= 1.0);  // also catches NaN (remaining norm 0, quirk A.6)
      s += fx_of(dbits(p <= 1.0 ? p : 0.0), 1023u);
    }
  }
  if (bad) atomicOr(&sc->pflag, 1u);
  u128 tot;
  thpre[(int64_t)blockIdx.x * NT + threadIdx.x] = block_excl_scan_u128<NT>(s, &tot, sm);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

// Exact prefix of FXP(p) before the first element of thread t of tile b and
// after its last one (tsum scanned: tile b's exclusive prefix at b, the total
// at nt).
__device__ inline void thread_range(const u128* __restrict__ tsum, const u128* __restrict__ thpre, int64_t b, int t,
                                    u128& lo, u128& hi) {
  const u128 t0 = tsum[b];
  const int64_t g = b * NT + t;
  lo = thpre[g] + t0;
  hi = t < NT - 1 ? thpre[g + 1] + t0 : tsum[b + 1];
}

__device__ inline int64_t first_k_at_least(double P, double u) {
  // smallest k >= 0 with fl(k + u) >= P (fl(k+u) is monotone in k); interval
  // ends are finite and below 2^32 by construction
  if (!(P < 4294967296.0)) return (int64_t)4294967296;
  double g = floor(P - u) - 2.0;
  int64_t k = g > 0.0 ? (int64_t)g : 0;
  while ((double)k + u < P) k++;
  return k;
}

// (T, tie) with  a < u128_to_double(pre, -96)  <=>  pre > T || (pre == T && tie)
// for a double a in [0, 2^32): the rounded prefix exceeds a exactly when the
// prefix lies above the midpoint between a and the next double, or on it when
// round-to-nearest-even goes up there (a's significand odd).  The midpoint in
// units of 2^-96 is (2m+1) 2^(E+95) for a = m 2^E; when it has a fractional
// part, "above" is "above its floor" and no tie can occur.
__device__ inline void above_threshold(double a, u128& T, bool& tie) {
  const uint64_t b = dbits(a);
  if (b == 0) {  // a == 0: any positive prefix rounds above it
    T = 0;
    tie = false;
    return;
  }
  int ex = (int)(b >> 52) & 0x7ff;
  uint64_t m = b & ((1ull << 52) - 1);
  if (ex == 0)
    ex = 1;
  else
    m |= 1ull << 52;
  const int sh = ex - 1075 + 95;  // a = m 2^(ex-1075); sh <= 74 for a < 2^32
  const u128 q = (u128)(2 * m + 1);
  if (sh >= 0) {
    T = q << sh;
    tie = (m & 1ull) != 0;
  } else {
    T = -sh >= 128 ? (u128)0 : q >> (-sh);
    tie = false;
  }
}

// Emission, one thread per ITEMS consecutive elements and no cooperation
// between threads: a thread's elements hold lattice points k+u exactly when
// its whole range [P(first), P(last+1)) does (first_k_at_least of the rounded
// prefix is monotone), so only those threads read their elements (one 128-byte
// line each) and walk them.  NS samples of one compression (uniforms in their
// scalars, everything else shared) are emitted in one pass: the prefixes,
// elements and probabilities are computed once, and each sample walks its own
// lattice, so each output is the one its single emission would write.
struct EmitOut {
  CompScalars* sc;
  uint32_t* idx;
  double* val;
};
template <int NS>
struct EmitOuts {
  EmitOut o[NS];
};

// Lattice points of the NS samples in the range of thread t of tile b:
// kc[q] the first, live[q] whether any lies in the range.
template <int NS>
__device__ inline bool range_points(const u128* __restrict__ tsum, const u128* __restrict__ thpre, int64_t b, int t,
                                    const double (&u)[NS], int64_t ns, u128& pre, int64_t (&kc)[NS],
                                    bool (&live)[NS]) {
  u128 end;
  thread_range(tsum, thpre, b, t, pre, end);
  const double Plo = u128_to_double(pre, -96), Phi = u128_to_double(end, -96);
  bool any = false;
#pragma unroll
  for (int q = 0; q < NS; q++) {
    live[q] = false;
    kc[q] = 0;
  }
  if (!(Phi > Plo)) return false;
#pragma unroll
  for (int q = 0; q < NS; q++) {
    kc[q] = first_k_at_least(Plo, u[q]);
    int64_t khi = first_k_at_least(Phi, u[q]);
    if (khi > ns) khi = ns;
    live[q] = kc[q] < khi;
    any |= live[q];
  }
  return any;
}

template <int NS>
__device__ inline void sample_uniforms(const EmitOuts<NS>& eo, const CompScalars& c, double (&u)[NS]) {
#pragma unroll
  for (int q = 0; q < NS; q++) u[q] = q == 0 ? c.u : eo.o[q].sc->u;
}

// Emission in two passes.  k_samp_live: every thread finds whether its range
// holds lattice points of any sample and, if so, appends its id to a list
// (one atomic per block; the list's order does not matter, each range writes
// only its own output positions).  k_samp_walk: one thread per listed range
// walks its elements.  The walk is the divergent, arithmetic-heavy part, and
// about a third of the ranges hold points: listing them first keeps every lane
// of a walking wave busy.
template <int NS>
__global__ __launch_bounds__(NT) void k_samp_live(const u128* __restrict__ tsum, const u128* __restrict__ thpre,
                                                  EmitOuts<NS> eo, uint32_t* __restrict__ list,
                                                  uint32_t* __restrict__ count) {
  __shared__ uint32_t sm[NT / WAVE + 1];
  __shared__ uint32_t base;
  const CompScalars c = *eo.o[0].sc;
  bool live_any = false;
  if (c.n_sampled != 0) {
    double u[NS];
    sample_uniforms(eo, c, u);
    u128 pre;
    int64_t kc[NS];
    bool live[NS];
    live_any = range_points<NS>(tsum, thpre, blockIdx.x, threadIdx.x, u, (int64_t)c.n_sampled, pre, kc, live);
  }
  uint32_t tot;
  const uint32_t ex = block_excl_scan_u32<NT>(live_any ? 1u : 0u, &tot, sm);
  if (threadIdx.x == 0) base = tot ? atomicAdd(count, tot) : 0u;
  __syncthreads();
  if (live_any) list[base + ex] = (uint32_t)(blockIdx.x * NT + threadIdx.x);
}

template <int NS>
__global__ __launch_bounds__(NT) void k_samp_walk(const double* __restrict__ x, const double* __restrict__ pin,
                                                  int64_t n, const u128* __restrict__ tsum,
                                                  const u128* __restrict__ thpre, EmitOuts<NS> eo,
                                                  const uint32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ count) {
  // every sample shares the first one's selection: n', K, remaining norm
  const CompScalars c = *eo.o[0].sc;
  if (c.n_sampled == 0) return;
  double u[NS];
  sample_uniforms(eo, c, u);
  const double nsd = (double)c.n_sampled;
  const int64_t ns = (int64_t)c.n_sampled, K = (int64_t)c.K;
  const uint32_t cnt = *count;
  const double* src = pin ? pin : x;
  for (uint32_t i = blockIdx.x * NT + threadIdx.x; i < cnt; i += gridDim.x * NT) {
    const uint32_t g = list[i];
    const int64_t b = g / NT;
    const int t = (int)(g % NT);
    u128 pre;
    int64_t kc[NS];
    bool live[NS];
    if (!range_points<NS>(tsum, thpre, b, t, u, ns, pre, kc, live)) continue;
    const int64_t base = b * TILE + (int64_t)t * ITEMS;
    double vals[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) vals[r] = base + r < n ? src[base + r] : 0.0;
    // Walk the lattice alongside the elements: kc = first k with fl(k+u) >= P
    // of the running prefix.  Element j owns the k with P_j <= fl(k+u) <
    // P_{j+1}, i.e. those kc passes while fl(kc+u) < P_{j+1} (the same set as
    // first_k_at_least(P_j) .. first_k_at_least(P_{j+1}), by monotonicity).
    // fl(kc+u) < P_{j+1} is tested on the exact prefix against the threshold
    // of fl(kc+u) (above_threshold), so the prefix is rounded to a double only
    // here
    u128 T[NS];
    bool tie[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) above_threshold((double)kc[q] + u[q], T[q], tie[q]);
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
      const int64_t j = base + r;
      const double v = vals[r];
      const double p = j < n ? (pin ? v : prob_of(v, (uint64_t)j, c, nsd)) : 0.0;
      const u128 f = fx_of(dbits(p <= 1.0 ? p : 0.0), 1023u);
      pre += f;
      if (f != 0) {
#pragma unroll
        for (int q = 0; q < NS; q++) {
          if (!live[q]) continue;
          int ne = 0;
          while (kc[q] < ns && (pre > T[q] || (pre == T[q] && tie[q]))) {
            eo.o[q].idx[K + kc[q]] = (uint32_t)j;
            if (eo.o[q].val) eo.o[q].val[K + kc[q]] = v / p;
            kc[q]++;
            ne++;
            above_threshold((double)kc[q] + u[q], T[q], tie[q]);
          }
          if (ne > 1) eo.o[q].sc->dup = 1u;  // j emitted more than once
        }
      }
    }
  }
}

__global__ void k_zero_count(uint32_t* count) { *count = 0u; }

// Scratch of the two-pass emission: a list of up to nt*NT range ids and its
// count (callers pass buffers that are idle during the emission).
struct EmitScratch {
  uint32_t* list;
  uint32_t* count;
};

// The emission of NS samples of one compression (or of given probabilities)
template <int NS>
static void samp_emit_n(const double* x, const double* pin, int64_t n, const EmitOuts<NS>& eo, const u128* tsum,
                        const u128* thpre, const EmitScratch& es, int64_t nt, hipStream_t s) {
  k_zero_count<<<1, 1, 0, s>>>(es.count);
  k_samp_live<NS><<<nt, NT, 0, s>>>(tsum, thpre, eo, es.list, es.count);
  const unsigned gw = (unsigned)std::min<int64_t>(nt, 4096);
  k_samp_walk<NS><<<gw, NT, 0, s>>>(x, pin, n, tsum, thpre, eo, es.list, es.count);
}

static void samp_emit(const double* x, const double* pin, int64_t n, CompScalars* sc, const u128* tsum,
                      const u128* thpre, uint32_t* out_idx, double* out_val, const EmitScratch& es, int64_t nt,
                      hipStream_t s) {
  EmitOuts<1> eo;
  eo.o[0] = EmitOut{sc, out_idx, out_val};
  samp_emit_n<1>(x, pin, n, eo, tsum, thpre, es, nt, s);
}


// Single thread: total interval end and the first not-found position.
__global__ void k_samp_final(CompScalars* sc, const u128* __restrict__ total) {
  const double u = sc->u;
  if (sc->n_sampled == 0) {
    sc->Ptot = 0.0;
    sc->k_notfound = 0;
    sc->n_notfound = 0;
    return;
  }
  double Ptot = u128_to_double(*total, -96);
  int64_t kn = first_k_at_least(Ptot, u);
  int64_t ns = (int64_t)sc->n_sampled;
  sc->Ptot = Ptot;
  sc->k_notfound = (uint64_t)(kn < ns ? kn : ns);
  sc->n_notfound = (uint64_t)(kn < ns ? ns - kn : 0);
  if (kn < ns) sc->dup = 1u;  // the not-found samples all take index 0
}

// Positions past the last interval keep index 0 (src/fri_utils.cpp:132).
__global__ void k_samp_notfound(const double* __restrict__ x, const CompScalars* __restrict__ sc,
                                uint32_t* __restrict__ out_idx, double* __restrict__ out_val) {
  const CompScalars c = *sc;
  const int64_t ns = (int64_t)c.n_sampled, K = (int64_t)c.K;
  const int64_t k = (int64_t)c.k_notfound + (int64_t)blockIdx.x * NT + threadIdx.x;
  if (k >= ns) return;
  out_idx[K + k] = 0;
  if (out_val) out_val[K + k] = x[0] / prob_of(x[0], 0, c, (double)c.n_sampled);
}

// ---------------------------------------------------------------- pivotal sampling
// sample_pivotal (src/fri_utils.cpp:63-127).  The chain carries a candidate i
// (initially 0) and a running mass a; step j (j >= 1, one uniform each)
// either completes a unit of mass or not.  Not completing: i = j with
// probability p_j / a.  Completing (a drops by 1): with probability
// (1 - p_j) / (1 - a) emit i and set i = j, else emit j.  The chain stops
// after the step whose emission makes n'; if it reaches the last step
// without that, a mass within 1e-5 of 1 emits the candidate.
//
// Canonical mass: Q_{j+1} = sum_{l<=j} FXP(p_l) exactly (units 2^-96); the
// units completed through step j are c_j = min(j, floor(Q_{j+1})) and
// a_j = round(Q_{j+1} - c_j).  So the mass at every step is a function of the
// prefix alone, and so is every acceptance probability: all steps decide in
// parallel.  The candidate before step j is the last accepted step before j
// (or 0) -- a max-scan over positions -- and a completing step j writes its
// emission at output position c_j - 1.  The stop step is the first j >= 1
// with c_j = n' (c is non-decreasing in steps of at most 1).
//
// Step j's uniform is U[j] when an array is given, else the j-th SplitMix64
// output seeded with the call's key (Steele, Lea and Flood, OOPSLA 2014).

__device__ inline double splitmix_u(uint64_t key, int64_t j) {
  uint64_t z = key + (uint64_t)(j + 1) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

// Units completed through step j (cn) and before it (cp), and the mass a_j,
// from the exact prefix before step j (Qp) and FXP(p_j) (F).
__device__ inline void piv_step(u128 Qp, u128 F, int64_t j, int64_t& cp, int64_t& cn, double& a) {
  const u128 Qn = Qp + F;
  const int64_t fn = (int64_t)(Qn >> 96), fp = (int64_t)(Qp >> 96);
  cn = fn < j ? fn : j;
  cp = j == 0 ? 0 : (fp < j - 1 ? fp : j - 1);
  a = u128_to_double(Qn - ((u128)cn << 96), -96);
}

__device__ inline bool piv_accept(double p, double a, bool cross, double uj) {
  return cross ? uj < (1.0 - p) / (1.0 - a) : uj < p / a;
}

// pst[0]: the stop step (~0 when none), pst[1]: the last accepted step overall.
__global__ void k_piv_reset(uint64_t* pst) {
  if (threadIdx.x == 0) {
    pst[0] = ~0ull;
    pst[1] = 0;
  }
}

// Per thread and per tile: the last accepted step (thmax, tmax); the stop step.
template <bool HASU>
__global__ __launch_bounds__(NT) void k_piv_mark(const double* __restrict__ x, const double* __restrict__ pin,
                                                 const double* __restrict__ U, int64_t n,
                                                 const CompScalars* __restrict__ sc, const u128* __restrict__ tsum,
                                                 const u128* __restrict__ thpre, uint32_t* __restrict__ tmax,
                                                 uint32_t* __restrict__ thmax, uint64_t* __restrict__ pst) {
  __shared__ uint32_t sm32[NT / WAVE + 1];
  __shared__ double lds[LTILE];
  __shared__ double ulds[HASU ? LTILE : 1];
  const CompScalars c = *sc;
  if (c.n_sampled == 0 && pin == nullptr) return;
  const double nsd = (double)c.n_sampled;
  const int64_t ns = (int64_t)c.n_sampled;
  const int64_t tb = (int64_t)blockIdx.x * TILE;
  const int64_t base = tb + (int64_t)threadIdx.x * ITEMS;
  u128 pre, pend;
  thread_range(tsum, thpre, blockIdx.x, threadIdx.x, pre, pend);
  stage_tile(pin ? pin : x, tb, n, lds);
  if (HASU) stage_tile(U, tb, n, ulds);
  uint32_t last = 0;
#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    const int64_t j = base + r;
    if (j < n) {
      const double v = tile_elem(lds, r);
      const double p = pin ? v : prob_of(v, (uint64_t)j, c, nsd);
      const u128 F = fx_of(dbits(p <= 1.0 ? p : 0.0), 1023u);
      if (j >= 1) {
        int64_t cp, cn;
        double a;
        piv_step(pre, F, j, cp, cn, a);
        const double uj = HASU ? tile_elem(ulds, r) : splitmix_u(c.key, j);
        if (piv_accept(p, a, cn > cp, uj)) last = (uint32_t)j;
        if (cn == ns && (j == 1 || cp < ns)) pst[0] = (uint64_t)j;  // the one stop step
      }
      pre += F;
    }
  }
  thmax[(int64_t)blockIdx.x * NT + threadIdx.x] = last;
  uint32_t bmax;
  (void)block_excl_max_u32<NT>(last, &bmax, sm32);
  if (threadIdx.x == 0) tmax[blockIdx.x] = bmax;
}

// Exclusive max-scan of the tile maxima in place (one block); pst[1] = overall.
constexpr int PSN = 1024, PSI = 8;
__global__ __launch_bounds__(PSN) void k_piv_scan(uint32_t* __restrict__ tmax, int64_t nt,
                                                  uint64_t* __restrict__ pst) {
  __shared__ uint32_t sm32[PSN / WAVE + 1];
  uint32_t carry = 0;
  for (int64_t b0 = 0; b0 < nt; b0 += (int64_t)PSN * PSI) {
    uint32_t v[PSI];
    uint32_t tm = 0;
#pragma unroll
    for (int k = 0; k < PSI; k++) {
      const int64_t i = b0 + (int64_t)threadIdx.x * PSI + k;
      v[k] = i < nt ? tmax[i] : 0u;
      tm = tm > v[k] ? tm : v[k];
    }
    uint32_t chunk;
    uint32_t run = block_excl_max_u32<PSN>(tm, &chunk, sm32);
    run = run > carry ? run : carry;
#pragma unroll
    for (int k = 0; k < PSI; k++) {
      const int64_t i = b0 + (int64_t)threadIdx.x * PSI + k;
      if (i < nt) tmax[i] = run;
      run = run > v[k] ? run : v[k];
    }
    carry = carry > chunk ? carry : chunk;
  }
  if (threadIdx.x == 0) pst[1] = carry;
}

// Emission: threads whose steps complete a unit (up to the stop step) walk
// them with the candidate carried in from the max-scan.
template <bool HASU>
__global__ __launch_bounds__(NT) void k_piv_emit(const double* __restrict__ x, const double* __restrict__ pin,
                                                 const double* __restrict__ U, int64_t n,
                                                 const CompScalars* __restrict__ sc, const u128* __restrict__ tsum,
                                                 const u128* __restrict__ thpre, const uint32_t* __restrict__ tmax,
                                                 const uint32_t* __restrict__ thmax, const uint64_t* __restrict__ pst,
                                                 uint32_t* __restrict__ out_idx, double* __restrict__ out_val) {
  __shared__ uint32_t sm32[NT / WAVE + 1];
  __shared__ double lds[LTILE];
  __shared__ double ulds[HASU ? LTILE : 1];
  const CompScalars c = *sc;
  if (c.n_sampled == 0 && pin == nullptr) return;
  const double nsd = (double)c.n_sampled;
  const int64_t ns = (int64_t)c.n_sampled, K = (int64_t)c.K;
  const uint64_t jb = pst[0];
  const int64_t jstop = jb < (uint64_t)n ? (int64_t)jb : n - 1;  // last step taken
  const int64_t tb = (int64_t)blockIdx.x * TILE;
  const int64_t base = tb + (int64_t)threadIdx.x * ITEMS;
  u128 pre, pend;
  thread_range(tsum, thpre, blockIdx.x, threadIdx.x, pre, pend);
  const u128 st = pend - pre;
  // does a step in [max(base,1), min(base+ITEMS-1, jstop)] complete a unit?
  bool work = false;
  {
    const int64_t j0 = base > 1 ? base : 1;
    const int64_t jl = base + ITEMS - 1 < n - 1 ? base + ITEMS - 1 : n - 1;
    const int64_t j1 = jl < jstop ? jl : jstop;
    if (j0 <= j1) {
      if (j1 < jl) {
        work = true;  // the stop step lies in this thread's range
      } else {
        const int64_t fb = (int64_t)(pre >> 96), fe = (int64_t)((pre + st) >> 96);
        const int64_t cb = base == 0 ? 0 : (fb < base - 1 ? fb : base - 1);
        const int64_t ce = fe < jl ? fe : jl;
        work = ce > cb;
      }
    }
  }
  if (!__syncthreads_or(work)) return;
  uint32_t bmax;
  uint32_t cand = block_excl_max_u32<NT>(thmax[(int64_t)blockIdx.x * NT + threadIdx.x], &bmax, sm32);
  cand = cand > tmax[blockIdx.x] ? cand : tmax[blockIdx.x];
  stage_tile(pin ? pin : x, tb, n, lds);
  if (HASU) stage_tile(U, tb, n, ulds);
  if (!work) return;
#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    const int64_t j = base + r;
    if (j < n && j <= jstop) {
      const double v = tile_elem(lds, r);
      const double p = pin ? v : prob_of(v, (uint64_t)j, c, nsd);
      const u128 F = fx_of(dbits(p <= 1.0 ? p : 0.0), 1023u);
      if (j >= 1) {
        int64_t cp, cn;
        double a;
        piv_step(pre, F, j, cp, cn, a);
        const double uj = HASU ? tile_elem(ulds, r) : splitmix_u(c.key, j);
        const bool cross = cn > cp;
        const bool acc = piv_accept(p, a, cross, uj);
        if (cross && cn - 1 < ns) {
          const uint32_t e = acc ? cand : (uint32_t)j;
          out_idx[K + cn - 1] = e;
          if (out_val) {
            const double xe = acc ? x[e] : v;
            out_val[K + cn - 1] = xe / (acc ? prob_of(xe, e, c, nsd) : p);
          }
        }
        if (acc) cand = (uint32_t)j;
      }
      pre += F;
    }
  }
}

// The end of the chain (one thread): with no stop step, the last step's mass
// within 1e-5 of 1 emits the candidate; any other count than n' is an exit of
// the reference (src/fri_utils.cpp:100-123).
__global__ void k_piv_final(const double* __restrict__ x, const double* __restrict__ pin, int64_t n,
                            CompScalars* sc, const u128* __restrict__ total, const uint64_t* __restrict__ pst,
                            uint32_t* __restrict__ out_idx, double* __restrict__ out_val) {
  const CompScalars c = *sc;
  if (c.n_sampled == 0 && pin == nullptr) return;
  const int64_t ns = (int64_t)c.n_sampled, K = (int64_t)c.K;
  if (pst[0] < (uint64_t)n) return;  // the chain stopped with n' emissions
  if (n == 1) {
    if (ns != 0) sc->pivbad = 1u;
    return;
  }
  const u128 Q = *total;
  const int64_t fq = (int64_t)(Q >> 96);
  const int64_t cl = fq < n - 1 ? fq : n - 1;
  const double a = u128_to_double(Q - ((u128)cl << 96), -96);
  if (!(fabs(a - 1.0) < 1e-5)) {
    sc->pivbad = 1u;
    return;
  }
  if (cl < ns) {
    const uint32_t e = (uint32_t)pst[1];
    out_idx[K + cl] = e;
    if (out_val) out_val[K + cl] = x[e] / prob_of(x[e], e, c, (double)ns);
  }
  if (cl + 1 != ns) sc->pivbad = 1u;
}

// Sample positions [K, cap) start as index 0, value 0: a chain that ends
// short of n' (reported as an error) leaves no position unwritten.
__global__ void k_piv_clear(const CompScalars* __restrict__ sc, int64_t cap, uint32_t* __restrict__ out_idx,
                            double* __restrict__ out_val) {
  const int64_t k = (int64_t)sc->K + (int64_t)blockIdx.x * NT + threadIdx.x;
  if (k >= cap) return;
  out_idx[k] = 0u;
  if (out_val) out_val[k] = 0.0;
}

// The pivotal passes after the probability prefix (k_samp_tilesums and its
// tile scan, which left ts exclusive and tot = the total); cap = K + n' at
// most (m, or n_sample for the free function).
static int pivotal_passes(const double* x, const double* pin, const double* U, int64_t n, int64_t cap,
                          CompScalars* sc, const u128* ts, const u128* ths, const u128* tot, uint32_t* tmax,
                          uint32_t* thmax, uint64_t* pst, uint32_t* out_idx, double* out_val, hipStream_t s) {
  const int64_t nt = ceil_div(n, TILE);
  if (cap > 0) k_piv_clear<<<ceil_div(cap, NT), NT, 0, s>>>(sc, cap, out_idx, out_val);
  k_piv_reset<<<1, 64, 0, s>>>(pst);
  if (U) {
    k_piv_mark<true><<<nt, NT, 0, s>>>(x, pin, U, n, sc, ts, ths, tmax, thmax, pst);
  } else {
    k_piv_mark<false><<<nt, NT, 0, s>>>(x, pin, U, n, sc, ts, ths, tmax, thmax, pst);
  }
  k_piv_scan<<<1, PSN, 0, s>>>(tmax, nt, pst);
  if (U) {
    k_piv_emit<true><<<nt, NT, 0, s>>>(x, pin, U, n, sc, ts, ths, tmax, thmax, pst, out_idx, out_val);
  } else {
    k_piv_emit<false><<<nt, NT, 0, s>>>(x, pin, U, n, sc, ts, ths, tmax, thmax, pst, out_idx, out_val);
  }
  k_piv_final<<<1, 1, 0, s>>>(x, pin, n, sc, tot, pst, out_idx, out_val);
  FRICC_HIP(hipGetLastError());
  return OK;
}

// ---------------------------------------------------------------- parallel_sample
// Thread-partitioned pivotal sampling (src/fri_utils.cpp:176-362) with T
// chunks for the reference's T OpenMP threads; chunk j = [j*step, (j+1)*step),
// the last one taking the remainder, step = n / T.  Canonical sums as in the
// pivotal chain: |q| and |q_j| are exact sums of FXP(q) rounded once; the move
// of a chunk toward its sample count keeps an exact running sum.  Draws by
// position: the chunk chain's step j takes position j*step, chunk j's chain
// its step i position j*step + i.

struct PsPlan {
  int64_t g, lb;  // samples of the chunk, its first output position
  double a, t;    // expected count n'|q_j|/|q| and its fractional part
  int32_t dir;    // move: +1 up (g > a), -1 down (g < a), 0 none
  int32_t pad;
};

// Exact sum of FXP(q) per chunk (one block each); q outside [0, 1] flagged.
__global__ __launch_bounds__(NT) void k_ps_chunksum(const double* __restrict__ q, int64_t n, int64_t T, int64_t step,
                                                    u128* __restrict__ QJ, CompScalars* sc) {
  __shared__ u128 sm[NT / WAVE + 1];
  const int64_t j = blockIdx.x;
  const int64_t lo = j * step, hi = j == T - 1 ? n : lo + step;
  u128 s = 0;
  uint32_t bad = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += NT) {
    const double v = q[i];
    bad |= !(v >= 0.0 && v <= 1.0);
    s += fx_of(dbits(v >= 0.0 && v <= 1.0 ? v : 0.0), 1023u);
  }
  if (bad) atomicOr(&sc->pflag, 1u);
  u128 tot;
  (void)block_excl_scan_u128<NT>(s, &tot, sm);
  if (threadIdx.x == 0) QJ[j] = tot;
}

// The plan (one thread): expected counts, the chunk chain over the
// fractional parts (canonical, serial: T is small), output offsets.
__global__ void k_ps_plan(const u128* __restrict__ QJ, int64_t T, int64_t step, const double* __restrict__ U,
                          PsPlan* __restrict__ plan, CompScalars* sc) {
  if (threadIdx.x != 0) return;
  const int64_t ns = (int64_t)sc->n_sampled;
  const uint64_t key = sc->key;
  u128 Qt = 0;
  for (int64_t j = 0; j < T; j++) Qt += QJ[j];
  const double qn = u128_to_double(Qt, -96);
  if (qn == 0.0) {
    for (int64_t j = 0; j < T; j++) plan[j] = PsPlan{0, 0, 0.0, 0.0, 0, 0};
    if (ns != 0) sc->pivbad = 1u;
    return;
  }
  int64_t c = ns;
  for (int64_t j = 0; j < T; j++) {
    const double a = (double)ns * u128_to_double(QJ[j], -96) / qn;
    const double fa = floor(a);
    plan[j] = PsPlan{(int64_t)fa, 0, a, a - fa, 0, 0};
    c -= (int64_t)fa;
  }
  if (c < 0) {
    sc->pivbad = 1u;
    return;
  }
  // the chunk chain over t (the canonical chain of sample_pivotal, serially)
  {
    int64_t cand = 0, k = 0, cprev = 0;
    u128 Q = fx_of(dbits(plan[0].t), 1023u);
    bool done = false;
    for (int64_t j = 1; j < T && !done; j++) {
      const double tj = plan[j].t;
      const double uj = U ? U[j * step] : splitmix_u(key, j * step);
      Q += fx_of(dbits(tj), 1023u);
      const int64_t fq = (int64_t)(Q >> 96);
      const int64_t cn = fq < j ? fq : j;
      const bool cross = cn > cprev;
      cprev = cn;
      const double a = u128_to_double(Q - ((u128)cn << 96), -96);
      if (!cross) {
        if (uj < tj / a) cand = j;
      } else {
        const double pa = (1.0 - tj) / (1.0 - a);
        if (pa > 1.0 || pa < 0.0) {
          sc->pivbad = 1u;
          return;
        }
        if (uj < pa) {
          if (k < c) plan[cand].g += 1;
          k++;
          cand = j;
        } else {
          if (k < c) plan[j].g += 1;
          k++;
        }
      }
      if (k == c) {
        done = true;
      } else if (j == T - 1) {
        if (fabs(a - 1.0) < 1e-5) {
          if (k < c) plan[cand].g += 1;
          k++;
        } else {
          sc->pivbad = 1u;
          return;
        }
      }
    }
    if (k != c) {
      sc->pivbad = 1u;
      return;
    }
  }
  int64_t lb = 0;
  for (int64_t j = 0; j < T; j++) {
    plan[j].lb = lb;
    lb += plan[j].g;
    const double gd = (double)plan[j].g;
    plan[j].dir = gd > plan[j].a ? 1 : (gd < plan[j].a ? -1 : 0);
  }
}

// Chunk range [lo, hi) staged tile by tile: the TILE elements from lo + tb.
__device__ inline void stage_range(const double* __restrict__ src, int64_t base, int64_t end, double* lds) {
#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    const int e = r * NT + threadIdx.x;
    const int64_t i = base + e;
    lds[e + e / ITEMS] = i < end ? src[i] : 0.0;
  }
  __syncthreads();
}

// The move of each chunk toward g (one block per chunk), in index order:
// up y = min(1, q/t), down y = max(0, (q - t)/(1 - t)); the running sum
// S_i = |q_j| + sum_{l<=i} (FXP(y_l) - FXP(q_l)) (exact, monotone) first
// reaching g (up: S >= g, down: S <= g) at i, which takes FXP(y_i) + g - S_i;
// later elements keep q.  Writes the chunk's values to qa.
__global__ __launch_bounds__(NT) void k_ps_move(const double* __restrict__ q, int64_t n, int64_t T, int64_t step,
                                                const u128* __restrict__ QJ, const PsPlan* __restrict__ plan,
                                                double* __restrict__ qa, CompScalars* sc) {
  __shared__ u128 sm[NT / WAVE + 1];
  __shared__ double lds[LTILE];
  __shared__ int cross_seen;
  const int64_t j = blockIdx.x;
  const PsPlan P = plan[j];
  if (P.dir == 0) return;
  const bool up = P.dir > 0;
  const int64_t lo = j * step, hi = j == T - 1 ? n : lo + step;
  const double tj = P.t;
  const u128 G = (u128)P.g << 96, Q0 = QJ[j];
  u128 D = 0;  // |sum of the changes so far|
  if (threadIdx.x == 0) cross_seen = 0;
  __syncthreads();
  for (int64_t tb = lo; tb < hi; tb += TILE) {
    stage_range(q, tb, hi, lds);
    const bool before = cross_seen == 0;
    const int64_t base = tb + (int64_t)threadIdx.x * ITEMS;
    u128 ds = 0;
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
      const int64_t i = base + r;
      if (i < hi) {
        const double v = tile_elem(lds, r);
        double y;
        if (up) {
          const double rr = v / tj;
          y = rr < 1.0 ? rr : 1.0;
        } else {
          const double rr = (v - tj) / (1.0 - tj);
          y = 0.0 < rr ? rr : 0.0;
        }
        const u128 Fy = fx_of(dbits(y), 1023u), Fq = fx_of(dbits(v), 1023u);
        ds += up ? Fy - Fq : Fq - Fy;
      }
    }
    u128 tot;
    u128 pre = block_excl_scan_u128<NT>(ds, &tot, sm) + D;
    bool found = false;
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
      const int64_t i = base + r;
      if (i < hi) {
        const double v = tile_elem(lds, r);
        double out = v;
        if (before) {
          double y;
          if (up) {
            const double rr = v / tj;
            y = rr < 1.0 ? rr : 1.0;
          } else {
            const double rr = (v - tj) / (1.0 - tj);
            y = 0.0 < rr ? rr : 0.0;
          }
          const u128 Fy = fx_of(dbits(y), 1023u), Fq = fx_of(dbits(v), 1023u);
          const u128 Sp = up ? Q0 + pre : Q0 - pre;
          pre += up ? Fy - Fq : Fq - Fy;
          const u128 S = up ? Q0 + pre : Q0 - pre;
          const bool hit = up ? S >= G : S <= G;
          const bool hitp = i == lo ? false : (up ? Sp >= G : Sp <= G);
          if (!hitp) {
            out = y;
            if (hit) {
              // negative (or above 1) only when the chunk held more than g
              // already (q not summing to n'): reported like any p outside [0, 1]
              out = Fy + G >= S ? u128_to_double(Fy + G - S, -96) : -u128_to_double(S - Fy - G, -96);
              if (!(out >= 0.0 && out <= 1.0)) atomicOr(&sc->pflag, 1u);
              found = true;
            }
          }
        }
        qa[i] = out;
      }
    }
    D += tot;
    if (__syncthreads_or(found) && threadIdx.x == 0) cross_seen = 1;
    __syncthreads();
  }
}

// Each chunk's own canonical pivotal chain (one block per chunk, its tiles in
// order): g samples written at plan.lb, indices shifted by the chunk start.
template <bool HASU>
__global__ __launch_bounds__(NT) void k_ps_chain(const double* __restrict__ q, const double* __restrict__ qa,
                                                 const double* __restrict__ U, int64_t n, int64_t T, int64_t step,
                                                 int64_t ns_total, const PsPlan* __restrict__ plan, CompScalars* sc,
                                                 uint32_t* __restrict__ out_idx) {
  __shared__ u128 sm[NT / WAVE + 1];
  __shared__ uint32_t sm32[NT / WAVE + 1];
  __shared__ double lds[LTILE];
  __shared__ double ulds[HASU ? LTILE : 1];
  __shared__ int stop_seen;
  const int64_t j = blockIdx.x;
  const PsPlan P = plan[j];
  const uint64_t key = sc->key;
  const int64_t lo = j * step, hi = j == T - 1 ? n : lo + step, len = hi - lo;
  const double* src = P.dir != 0 ? qa : q;
  const int64_t g = P.g;
  u128 Q = 0;          // exact prefix before the tile
  uint32_t cenc = 0;   // candidate so far, local index + 1 (0: the initial candidate, local 0)
  uint32_t bad = 0;
  if (threadIdx.x == 0) stop_seen = 0;
  __syncthreads();
  for (int64_t tb = 0; tb < len; tb += TILE) {
    stage_range(src, lo + tb, hi, lds);
    if (HASU) stage_range(U, lo + tb, hi, ulds);
    const int64_t base = tb + (int64_t)threadIdx.x * ITEMS;  // local index of the thread's first element
    u128 st = 0;
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
      const double v = tile_elem(lds, r);
      if (base + r < len) {
        bad |= !(v >= 0.0 && v <= 1.0);
        st += fx_of(dbits(v >= 0.0 && v <= 1.0 ? v : 0.0), 1023u);
      }
    }
    u128 tot;
    const u128 pre = block_excl_scan_u128<NT>(st, &tot, sm) + Q;
    // decisions: the last accepted step of this thread, and the stop step
    uint32_t last = 0;
    bool stop = false;
    {
      u128 Qp = pre;
#pragma unroll
      for (int r = 0; r < ITEMS; r++) {
        const int64_t li = base + r;
        if (li < len) {
          const double v = tile_elem(lds, r);
          const double p = v >= 0.0 && v <= 1.0 ? v : 0.0;
          const u128 F = fx_of(dbits(p), 1023u);
          if (li >= 1) {
            int64_t cp, cn;
            double a;
            piv_step(Qp, F, li, cp, cn, a);
            const double uj = HASU ? tile_elem(ulds, r) : splitmix_u(key, lo + li);
            if (piv_accept(v, a, cn > cp, uj)) last = (uint32_t)(li + 1);
            if (cn == g && (li == 1 || cp < g)) stop = true;
          }
          Qp += F;
        }
      }
    }
    uint32_t bmax;
    uint32_t cin = block_excl_max_u32<NT>(last, &bmax, sm32);
    cin = cin > cenc ? cin : cenc;
    // emissions at completing steps
    {
      u128 Qp = pre;
      uint32_t cand = cin;
#pragma unroll
      for (int r = 0; r < ITEMS; r++) {
        const int64_t li = base + r;
        if (li < len) {
          const double v = tile_elem(lds, r);
          const double p = v >= 0.0 && v <= 1.0 ? v : 0.0;
          const u128 F = fx_of(dbits(p), 1023u);
          if (li >= 1) {
            int64_t cp, cn;
            double a;
            piv_step(Qp, F, li, cp, cn, a);
            const double uj = HASU ? tile_elem(ulds, r) : splitmix_u(key, lo + li);
            const bool cross = cn > cp;
            const bool acc = piv_accept(v, a, cross, uj);
            if (cross && cn - 1 < g && P.lb + cn - 1 < ns_total) {
              const int64_t e = acc ? (cand == 0 ? 0 : (int64_t)cand - 1) : li;
              out_idx[P.lb + cn - 1] = (uint32_t)(lo + e);
            }
            if (acc) cand = (uint32_t)(li + 1);
          }
          Qp += F;
        }
      }
    }
    Q += tot;
    cenc = cenc > bmax ? cenc : bmax;
    if (__syncthreads_or(stop) && threadIdx.x == 0) stop_seen = 1;
    __syncthreads();
    if (stop_seen) break;
  }
  if (__syncthreads_or(bad)) {
    if (threadIdx.x == 0) atomicOr(&sc->pflag, 1u);
    return;
  }
  if (threadIdx.x != 0 || stop_seen) return;
  // no stop step: the last step's mass within 1e-5 of 1 emits the candidate
  if (len == 1) {
    if (g != 0) sc->pivbad = 1u;
    return;
  }
  const int64_t fq = (int64_t)(Q >> 96);
  const int64_t cl = fq < len - 1 ? fq : len - 1;
  const double a = u128_to_double(Q - ((u128)cl << 96), -96);
  if (!(fabs(a - 1.0) < 1e-5)) {
    sc->pivbad = 1u;
    return;
  }
  if (cl < g && P.lb + cl < ns_total) out_idx[P.lb + cl] = (uint32_t)(lo + (cenc == 0 ? 0 : (int64_t)cenc - 1));
  if (cl + 1 != g) sc->pivbad = 1u;
}

// D part of the output (or all of it for "largest").
__global__ void k_emit_sorted(const double* __restrict__ x, const uint32_t* __restrict__ vals,
                              const CompScalars* __restrict__ sc, int64_t cnt_if_largest,
                              uint32_t* __restrict__ out_idx, double* __restrict__ out_val) {
  const int64_t cnt = cnt_if_largest >= 0 ? cnt_if_largest : (int64_t)sc->K;
  const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (k >= cnt) return;
  uint32_t j = vals[k];
  out_idx[k] = j;
  out_val[k] = x[j];
}

// Scalars of a call (u from u_dev when given, else u_val) and, when hist is
// given, the zeroed top-bit and digit histograms (a kernel rather than a
// memset, so the call can be recorded in a graph).
__global__ void k_init_scalars(CompScalars* sc, int64_t m, int64_t n_sampled, double u_val,
                               const double* __restrict__ u_dev, uint32_t* __restrict__ hist, int nhist) {
  if (threadIdx.x == 0) {
    CompScalars z = {};
    z.K = (unsigned long long)m;
    z.n_sampled = (uint64_t)n_sampled;
    const double u = u_dev ? *u_dev : u_val;
    z.ubad = (u >= 0.0 && u <= 1.0) ? 0u : 1u;
    z.u = z.ubad ? 0.0 : u;  // an out-of-range u is reported, never used
    z.key = dbits(z.u);
    *sc = z;
  }
  if (hist)
    for (int i = threadIdx.x; i < nhist; i += blockDim.x) hist[i] = 0u;
}

// Copy of the call's scalars for a later host read (compression statistics,
// error flags) without a round trip now.
__global__ void k_export_scalars(const CompScalars* __restrict__ sc, CompScalars* __restrict__ out) {
  if (threadIdx.x == 0) *out = *sc;
}

// ---------------------------------------------------------------- driver

size_t comp_scalars_size() { return sizeof(CompScalars); }

const uint32_t* comp_scalars_dup_flag(const void* dev_sc) { return &((const CompScalars*)dev_sc)->dup; }

// Host view of a call's scalars: statistics into info, and the call's status.
int comp_scalars_status(const void* host_sc, CompressInfo* info) {
  const CompScalars& h = *(const CompScalars*)host_sc;
  if (info) {
    memset(info, 0, sizeof(*info));
    info->n_D = (int64_t)h.K;
    info->n_sampled = (int64_t)h.n_sampled;
    info->n_notfound = (int64_t)h.n_notfound;
    info->norm0 = h.R0;
    info->remaining_norm = h.RK;
    info->p_total = h.Ptot;
    info->n_candidates = (int64_t)h.n_cand;
    info->nonfinite = (int32_t)h.nonfinite;
  }
  if (h.ubad) {
    set_error("compress: uniform must lie in [0, 1]");
    return EINVAL_;
  }
  if (h.nonfinite) {
    set_error("compress: non-finite value in the input vector");
    return ENONFINITE;
  }
  if (h.pflag) {
    set_error("compress: a sampling probability exceeds 1 (reference exits, src/fri_utils.cpp:383-395)");
    return EPROB;
  }
  if (h.pivbad) {
    set_error("compress: pivotal sampling ended with a sample count other than n' (reference exits, "
              "src/fri_utils.cpp:100-123)");
    return ESAMPLE;
  }
  return OK;
}

// Scratch of one compression of n elements.
struct CompScratch {
  CompScalars* sc;
  uint32_t* hist;   // 4096 top-bit buckets, then 256 digit totals
  uint32_t* tcnt;   // nt + 1 per-tile candidate counts / offsets
  uint32_t* ssum;   // tile sums of the u32 scan of tcnt
  u128* tot;        // total of the probability prefix
  uint64_t* kA;     // candidate keys and indices, two buffers each (radix ping-pong)
  uint32_t* vA;
  uint64_t* kB;
  uint32_t* vB;
  uint32_t* bh;     // per-tile digit histograms [256][nt]
  u128* ts;         // nt + 1 tile sums
  u128* ths;        // nt * NT per-thread sums
  uint32_t* tmax;   // pivotal: nt last accepted steps per tile (then their exclusive max-scan)
  uint32_t* thmax;  // pivotal: nt * NT last accepted steps per thread
  uint64_t* pst;    // pivotal: stop step, last accepted step overall
};

// Byte offsets of the parts of CompScratch inside one buffer (256-byte
// aligned parts), and the total.
struct CompLayout {
  size_t off[15];
  size_t total;
};

static CompLayout comp_layout(int64_t n) {
  const int64_t nt = ceil_div(n, TILE);
  const size_t sz[15] = {sizeof(CompScalars),
                         (size_t)(4096 + 256) * 4,
                         (size_t)(nt + 1) * 4,
                         (size_t)scan_u32_sums_words(nt + 1) * 4,
                         64,
                         (size_t)n * 8,
                         (size_t)n * 4,
                         (size_t)n * 8,
                         (size_t)std::max<int64_t>(n, nt * NT) * 4,  // also the emission's range list
                         (size_t)nt * 256 * 4,
                         (size_t)(nt + 1) * sizeof(u128),
                         (size_t)nt * NT * sizeof(u128),
                         (size_t)(nt + 1) * 4,
                         (size_t)nt * NT * 4,
                         16};
  CompLayout L;
  size_t o = 0;
  for (int i = 0; i < 15; i++) {
    L.off[i] = o;
    o += (sz[i] + 255) / 256 * 256;
  }
  L.total = o;
  return L;
}

static CompScratch comp_scratch_at(uint8_t* base, const CompLayout& L) {
  CompScratch w;
  w.sc = (CompScalars*)(base + L.off[0]);
  w.hist = (uint32_t*)(base + L.off[1]);
  w.tcnt = (uint32_t*)(base + L.off[2]);
  w.ssum = (uint32_t*)(base + L.off[3]);
  w.tot = (u128*)(base + L.off[4]);
  w.kA = (uint64_t*)(base + L.off[5]);
  w.vA = (uint32_t*)(base + L.off[6]);
  w.kB = (uint64_t*)(base + L.off[7]);
  w.vB = (uint32_t*)(base + L.off[8]);
  w.bh = (uint32_t*)(base + L.off[9]);
  w.ts = (u128*)(base + L.off[10]);
  w.ths = (u128*)(base + L.off[11]);
  w.tmax = (uint32_t*)(base + L.off[12]);
  w.thmax = (uint32_t*)(base + L.off[13]);
  w.pst = (uint64_t*)(base + L.off[14]);
  return w;
}

// Workspace slots of a compression (the resample reads the first sample's
// scalars and sums back from the same slots).
enum { WS_SC = 0, WS_HIST = 1, WS_TCNT = 2, WS_K0 = 3, WS_V0 = 4, WS_K1 = 5, WS_V1 = 6, WS_SH = 7,
       WS_BH = 8, WS_TS = 9, WS_TOT = 10, WS_THS = 11, WS_SC2 = 12, WS_SSUM = 13, WS_TMAX = 14,
       WS_THMAX = 15, WS_PST = 16 };

// The pivotal passes' scratch (slots of the calling thread's arena).
static int piv_scratch_ws(int64_t n, uint32_t** tmax, uint32_t** thmax, uint64_t** pst) {
  const int64_t nt = ceil_div(n, TILE);
  void *a, *b, *c;
  FRICC_TRY(ws_get(WS_TMAX, (size_t)(nt + 1) * 4, &a));
  FRICC_TRY(ws_get(WS_THMAX, (size_t)nt * NT * 4, &b));
  FRICC_TRY(ws_get(WS_PST, 16, &c));
  *tmax = (uint32_t*)a;
  *thmax = (uint32_t*)b;
  *pst = (uint64_t*)c;
  return OK;
}

static int comp_scratch_ws(int64_t n, CompScratch* w) {
  const int64_t nt = ceil_div(n, TILE);
  void *p_sc, *p_hist, *p_tcnt, *p_ssum, *p_k0, *p_v0, *p_k1, *p_v1, *p_sh, *p_tot, *p_ts, *p_ths;
  FRICC_TRY(ws_get(WS_SC, sizeof(CompScalars), &p_sc));
  FRICC_TRY(ws_get(WS_HIST, (4096 + 256) * 4, &p_hist));
  FRICC_TRY(ws_get(WS_TCNT, (size_t)(nt + 1) * 4, &p_tcnt));
  FRICC_TRY(ws_get(WS_SSUM, (size_t)scan_u32_sums_words(nt + 1) * 4, &p_ssum));
  FRICC_TRY(ws_get(WS_TOT, 64, &p_tot));
  FRICC_TRY(ws_get(WS_K0, (size_t)n * 8, &p_k0));
  FRICC_TRY(ws_get(WS_V0, (size_t)n * 4, &p_v0));
  FRICC_TRY(ws_get(WS_K1, (size_t)n * 8, &p_k1));
  FRICC_TRY(ws_get(WS_V1, (size_t)std::max<int64_t>(n, nt * NT) * 4, &p_v1));  // also the emission's range list
  FRICC_TRY(ws_get(WS_SH, (size_t)nt * 256 * 4, &p_sh));
  FRICC_TRY(ws_get(WS_TS, (size_t)(nt + 1) * sizeof(u128), &p_ts));
  FRICC_TRY(ws_get(WS_THS, (size_t)nt * NT * sizeof(u128), &p_ths));
  w->sc = (CompScalars*)p_sc;
  w->hist = (uint32_t*)p_hist;
  w->tcnt = (uint32_t*)p_tcnt;
  w->ssum = (uint32_t*)p_ssum;
  w->tot = (u128*)p_tot;
  w->kA = (uint64_t*)p_k0;
  w->vA = (uint32_t*)p_v0;
  w->kB = (uint64_t*)p_k1;
  w->vB = (uint32_t*)p_v1;
  w->bh = (uint32_t*)p_sh;
  w->ts = (u128*)p_ts;
  w->ths = (u128*)p_ths;
  return piv_scratch_ws(n, &w->tmax, &w->thmax, &w->pst);
}

// Emission scratch from the workspace arena: the second radix-sort value
// buffer and the first word of the histogram, both idle once the sort is done
// (sized here for vectors of n elements; the compression sized them already).
static int emit_scratch_ws(int64_t n, EmitScratch* es) {
  const int64_t nt = ceil_div(n, TILE);
  void *p_list, *p_cnt;
  FRICC_TRY(ws_get(WS_V1, (size_t)std::max<int64_t>(n, nt * NT) * 4, &p_list));
  FRICC_TRY(ws_get(WS_HIST, (4096 + 256) * 4, &p_cnt));
  es->list = (uint32_t*)p_list;
  es->count = (uint32_t*)p_cnt;
  return OK;
}

static int check_compress_args(int64_t n, int64_t m, int method, double u_val, const double* u_dev) {
  FRICC_CHECK(n > 0 && n < (int64_t)1 << 31, "compress: vector length must be in [1, 2^31)");
  FRICC_CHECK(m > 0 && m <= n, "compress: need 0 < m <= n (fricc/friccsd.py:68-72)");
  FRICC_CHECK(method == COMP_LARGEST || method == COMP_FRI_SYSTEMATIC || method == COMP_FRI_PIVOTAL,
              "compress: unknown compression method");
  FRICC_CHECK(u_dev || (u_val >= 0.0 && u_val <= 1.0), "compress: uniform must lie in [0, 1]");
  return OK;
}

// The whole compression on stream s with no host round trip, in the given
// scratch: every size the host does not know (candidate count, |D|, samples)
// stays on the device and the kernels read it there; buffers are sized for
// the worst case.  The call's scalars end in w.sc and, when export_sc is
// given, in that device buffer.  u comes from u_dev when given, else u_val.
// defer_emit (systematic FRI only): stop after the probability prefix; the
// caller runs the emission (compress_pair_launch emits two samples at once).
static int compress_pipeline(const double* x, int64_t n, int64_t m, int method, double u_val, const double* u_dev,
                             uint32_t* out_idx, double* out_val, void* export_sc, hipStream_t s,
                             const CompScratch& w, bool defer_emit = false) {
  const int64_t nt = ceil_div(n, TILE);
  CompScalars* sc = w.sc;
  uint32_t* hist = w.hist;
  uint32_t* dtot = hist + 4096;
  uint32_t* tcnt = w.tcnt;
  uint64_t* kA = w.kA;
  uint32_t* vA = w.vA;
  uint64_t* kB = w.kB;
  uint32_t* vB = w.vB;
  uint32_t* bh = w.bh;
  u128* ts = w.ts;

  k_init_scalars<<<1, NT, 0, s>>>(sc, m, 0, u_val, u_dev, hist, 4096);
  const int gA = (int)std::min<int64_t>(2048, ceil_div(n, NT * 8));
  k_hist_top<<<gA, NT, 0, s>>>(x, n, hist);
  k_select_top<<<1, NT, 0, s>>>(hist, m, sc);
  k_norm_count<<<nt, NT, 0, s>>>(x, n, sc, tcnt);
  FRICC_TRY(scan_u32(tcnt, tcnt, nt, nullptr, s, w.ssum));
  k_compact<<<nt, NT, 0, s>>>(x, n, sc, tcnt, kA, vA);
  // --- sort candidates (stable LSD, 8 byte digits)
  const int gS = (int)std::min<int64_t>(nt, 2048);
  for (int d = 0; d < 8; d++) {
    k_sort_count<<<gS, NT, 0, s>>>(kA, sc, 8 * d, bh, nt);
    k_digit_total<<<256, NT, 0, s>>>(bh, nt, sc, dtot);
    k_digit_offsets<<<256, NT, 0, s>>>(bh, nt, sc, dtot);
    k_sort_scatter<<<gS, NT, 0, s>>>(kA, vA, kB, vB, sc, 8 * d, bh, nt);
    std::swap(kA, kB);
    std::swap(vA, vB);
  }
  FRICC_HIP(hipGetLastError());
  if (method == COMP_LARGEST) {
    k_emit_sorted<<<ceil_div(m, NT), NT, 0, s>>>(x, vA, sc, m, out_idx, out_val);
  } else {
    // --- D test over the first m sorted candidates
    const int64_t ntm = ceil_div(m, TILE);
    k_dfx_tilesums<<<ntm, NT, 0, s>>>(kA, m, sc, ts);
    k_scan_u128_sums<<<1, 1024, 0, s>>>(ts, ntm, nullptr);
    k_dtest<<<ntm, NT, 0, s>>>(kA, m, sc, ts);
    k_dfinal<<<1, NT, 0, s>>>(kA, vA, m, sc, ts);
    k_emit_sorted<<<ceil_div(m, NT), NT, 0, s>>>(x, vA, sc, -1, out_idx, out_val);
    // --- sampling of the remainder: exact prefix of FXP(p), then emission
    k_samp_tilesums<<<nt, NT, 0, s>>>(x, nullptr, n, sc, ts, w.ths);
    k_scan_u128_sums<<<1, 1024, 0, s>>>(ts, nt, w.tot);
    if (method == COMP_FRI_PIVOTAL) {
      FRICC_TRY(pivotal_passes(x, nullptr, nullptr, n, m, sc, ts, w.ths, w.tot, w.tmax, w.thmax, w.pst, out_idx,
                               out_val, s));
    } else if (defer_emit) {
      FRICC_HIP(hipGetLastError());
      return OK;
    } else {
      samp_emit(x, nullptr, n, sc, ts, w.ths, out_idx, out_val, EmitScratch{w.vB, w.hist}, nt, s);
      k_samp_final<<<1, 1, 0, s>>>(sc, w.tot);
      k_samp_notfound<<<ceil_div(m, NT), NT, 0, s>>>(x, sc, out_idx, out_val);
    }
  }
  if (export_sc) k_export_scalars<<<1, 64, 0, s>>>(sc, (CompScalars*)export_sc);
  FRICC_HIP(hipGetLastError());
  return OK;
}

int compress_launch(const double* x, int64_t n, int64_t m, int method, double u_val, const double* u_dev,
                    uint32_t* out_idx, double* out_val, void* export_sc, hipStream_t s, void** sc_out) {
  FRICC_TRY(check_compress_args(n, m, method, u_val, u_dev));
  CompScratch w;
  FRICC_TRY(comp_scratch_ws(n, &w));
  *sc_out = w.sc;
  return compress_pipeline(x, n, m, method, u_val, u_dev, out_idx, out_val, export_sc, s, w);
}

// Scalars of a second sample of the same compression: everything but the
// uniform (and what the emission derives from it) is the first one's.
__global__ void k_resample_scalars(const CompScalars* __restrict__ src, const double* __restrict__ u_dev,
                                   CompScalars* __restrict__ sc) {
  if (threadIdx.x != 0) return;
  CompScalars c = *src;
  const double u = *u_dev;
  c.ubad = (u >= 0.0 && u <= 1.0) ? 0u : 1u;
  c.u = c.ubad ? 0.0 : u;
  c.key = dbits(c.u);
  c.pivbad = 0u;
  c.dup = 0u;
  c.Ptot = 0.0;
  c.k_notfound = 0;
  c.n_notfound = 0;
  *sc = c;
}

// The D part of the first sample (|D| read on the device).
__global__ void k_copy_kept(const uint32_t* __restrict__ idx, const double* __restrict__ val,
                            const CompScalars* __restrict__ sc, int64_t m, uint32_t* __restrict__ out_idx,
                            double* __restrict__ out_val) {
  const int64_t K = (int64_t)sc->K;
  for (int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x; k < K && k < m; k += (int64_t)gridDim.x * NT) {
    out_idx[k] = idx[k];
    out_val[k] = val[k];
  }
}

// A second FRI sample of the vector compress_launch has just compressed (same
// x, n, m, method, stream and workspace arena, nothing in between): the
// selection, |D|, the remaining norm and the exact prefix sums of p do not
// depend on the uniform, so only the systematic emission runs again, with
// u from u_dev.  Its output is bit for bit that of a full compression with
// that uniform.  idx/val: the first sample's output (its D part is copied).
int compress_resample(const double* x, int64_t n, int64_t m, int method, const double* u_dev, const uint32_t* idx,
                      const double* val, uint32_t* out_idx, double* out_val, void* export_sc, hipStream_t s) {
  FRICC_CHECK(u_dev != nullptr, "compress: resampling needs a device uniform");
  const int64_t nt = ceil_div(n, TILE);
  void *p_sc, *p_sc2, *p_tot, *p_ts, *p_ths;
  FRICC_TRY(ws_get(WS_SC, sizeof(CompScalars), &p_sc));
  FRICC_TRY(ws_get(WS_SC2, sizeof(CompScalars), &p_sc2));
  FRICC_TRY(ws_get(WS_TOT, 64, &p_tot));
  FRICC_TRY(ws_get(WS_TS, (size_t)(nt + 1) * sizeof(u128), &p_ts));
  FRICC_TRY(ws_get(WS_THS, (size_t)nt * NT * sizeof(u128), &p_ths));
  CompScalars* sc = (CompScalars*)p_sc2;
  k_resample_scalars<<<1, 64, 0, s>>>((const CompScalars*)p_sc, u_dev, sc);
  // the kept part (all m entries for "largest", where K stays m: the second
  // sample is the first)
  k_copy_kept<<<(unsigned)std::min<int64_t>(ceil_div(m, NT), 2048), NT, 0, s>>>(idx, val, sc, m, out_idx, out_val);
  if (method == COMP_FRI_PIVOTAL) {
    uint32_t *tmax, *thmax;
    uint64_t* pst;
    FRICC_TRY(piv_scratch_ws(n, &tmax, &thmax, &pst));
    FRICC_TRY(pivotal_passes(x, nullptr, nullptr, n, m, sc, (const u128*)p_ts, (const u128*)p_ths,
                             (const u128*)p_tot, tmax, thmax, pst, out_idx, out_val, s));
  } else if (method != COMP_LARGEST) {
    EmitScratch es;
    FRICC_TRY(emit_scratch_ws(n, &es));
    samp_emit(x, nullptr, n, sc, (const u128*)p_ts, (const u128*)p_ths, out_idx, out_val, es, nt, s);
    k_samp_final<<<1, 1, 0, s>>>(sc, (u128*)p_tot);
    k_samp_notfound<<<ceil_div(m, NT), NT, 0, s>>>(x, sc, out_idx, out_val);
  }
  if (export_sc) k_export_scalars<<<1, 64, 0, s>>>(sc, (CompScalars*)export_sc);
  FRICC_HIP(hipGetLastError());
  return OK;
}

// Two samples of one compression (uniforms u_dev, u_dev2), bit for bit
// compress_launch with u_dev followed by compress_resample with u_dev2.  For
// systematic FRI the selection and the probability prefix run once and one
// pass emits both samples (samp_emit_n<2>); the other methods take the two
// calls.
int compress_pair_launch(const double* x, int64_t n, int64_t m, int method, const double* u_dev,
                         const double* u_dev2, uint32_t* idx1, double* val1, void* export1, uint32_t* idx2,
                         double* val2, void* export2, hipStream_t s) {
  FRICC_CHECK(u_dev != nullptr && u_dev2 != nullptr, "compress: a pair needs two device uniforms");
  if (method != COMP_FRI_SYSTEMATIC) {
    void* sc = nullptr;
    FRICC_TRY(compress_launch(x, n, m, method, 0.0, u_dev, idx1, val1, export1, s, &sc));
    return compress_resample(x, n, m, method, u_dev2, idx1, val1, idx2, val2, export2, s);
  }
  FRICC_TRY(check_compress_args(n, m, method, 0.0, u_dev));
  CompScratch w;
  FRICC_TRY(comp_scratch_ws(n, &w));
  FRICC_TRY(compress_pipeline(x, n, m, method, 0.0, u_dev, idx1, val1, nullptr, s, w, true));
  void* p_sc2;
  FRICC_TRY(ws_get(WS_SC2, sizeof(CompScalars), &p_sc2));
  CompScalars* sc2 = (CompScalars*)p_sc2;
  const int64_t nt = ceil_div(n, TILE);
  // the second sample's scalars (taken before the first emission: it changes
  // only what k_resample_scalars resets) and its kept part
  k_resample_scalars<<<1, 64, 0, s>>>(w.sc, u_dev2, sc2);
  k_copy_kept<<<(unsigned)std::min<int64_t>(ceil_div(m, NT), 2048), NT, 0, s>>>(idx1, val1, sc2, m, idx2, val2);
  EmitOuts<2> eo;
  eo.o[0] = EmitOut{w.sc, idx1, val1};
  eo.o[1] = EmitOut{sc2, idx2, val2};
  samp_emit_n<2>(x, nullptr, n, eo, w.ts, w.ths, EmitScratch{w.vB, w.hist}, nt, s);
  k_samp_final<<<1, 1, 0, s>>>(w.sc, w.tot);
  k_samp_notfound<<<ceil_div(m, NT), NT, 0, s>>>(x, w.sc, idx1, val1);
  k_samp_final<<<1, 1, 0, s>>>(sc2, w.tot);
  k_samp_notfound<<<ceil_div(m, NT), NT, 0, s>>>(x, sc2, idx2, val2);
  if (export1) k_export_scalars<<<1, 64, 0, s>>>(w.sc, (CompScalars*)export1);
  if (export2) k_export_scalars<<<1, 64, 0, s>>>(sc2, (CompScalars*)export2);
  FRICC_HIP(hipGetLastError());
  return OK;
}

// Samples a and b (uniforms u_a, u_b) of the vector compress_launch or
// compress_pair_launch has just compressed (same x, n, m, method, stream and
// arena, nothing in between): bit for bit two compress_resample calls.  For
// systematic FRI one pass emits both (samp_emit_n<2>).  idx/val: the first
// sample's output (its kept part is copied).
int compress_resample_pair(const double* x, int64_t n, int64_t m, int method, const double* u_a, const double* u_b,
                           const uint32_t* idx, const double* val, uint32_t* idx_a, double* val_a, void* export_a,
                           uint32_t* idx_b, double* val_b, void* export_b, hipStream_t s) {
  FRICC_CHECK(u_a != nullptr && u_b != nullptr, "compress: resampling needs device uniforms");
  if (method != COMP_FRI_SYSTEMATIC) {
    FRICC_TRY(compress_resample(x, n, m, method, u_a, idx, val, idx_a, val_a, export_a, s));
    return compress_resample(x, n, m, method, u_b, idx, val, idx_b, val_b, export_b, s);
  }
  const int64_t nt = ceil_div(n, TILE);
  void *p_sc, *p_sc2, *p_tot, *p_ts, *p_ths;
  FRICC_TRY(ws_get(WS_SC, sizeof(CompScalars), &p_sc));
  FRICC_TRY(ws_get(WS_SC2, 2 * sizeof(CompScalars), &p_sc2));
  FRICC_TRY(ws_get(WS_TOT, 64, &p_tot));
  FRICC_TRY(ws_get(WS_TS, (size_t)(nt + 1) * sizeof(u128), &p_ts));
  FRICC_TRY(ws_get(WS_THS, (size_t)nt * NT * sizeof(u128), &p_ths));
  CompScalars* sa = (CompScalars*)p_sc2;
  CompScalars* sb = sa + 1;
  const unsigned gk = (unsigned)std::min<int64_t>(ceil_div(m, NT), 2048);
  k_resample_scalars<<<1, 64, 0, s>>>((const CompScalars*)p_sc, u_a, sa);
  k_resample_scalars<<<1, 64, 0, s>>>((const CompScalars*)p_sc, u_b, sb);
  k_copy_kept<<<gk, NT, 0, s>>>(idx, val, sa, m, idx_a, val_a);
  k_copy_kept<<<gk, NT, 0, s>>>(idx, val, sb, m, idx_b, val_b);
  EmitOuts<2> eo;
  eo.o[0] = EmitOut{sa, idx_a, val_a};
  eo.o[1] = EmitOut{sb, idx_b, val_b};
  EmitScratch es;
  FRICC_TRY(emit_scratch_ws(n, &es));
  samp_emit_n<2>(x, nullptr, n, eo, (const u128*)p_ts, (const u128*)p_ths, es, nt, s);
  for (CompScalars* sc : {sa, sb}) k_samp_final<<<1, 1, 0, s>>>(sc, (u128*)p_tot);
  k_samp_notfound<<<ceil_div(m, NT), NT, 0, s>>>(x, sa, idx_a, val_a);
  k_samp_notfound<<<ceil_div(m, NT), NT, 0, s>>>(x, sb, idx_b, val_b);
  if (export_a) k_export_scalars<<<1, 64, 0, s>>>(sa, (CompScalars*)export_a);
  if (export_b) k_export_scalars<<<1, 64, 0, s>>>(sb, (CompScalars*)export_b);
  FRICC_HIP(hipGetLastError());
  return OK;
}

int compress(const double* x, int64_t n, int64_t m, int method, double u, uint32_t* out_idx, double* out_val,
             CompressInfo* info, hipStream_t s, int* dup) {
  memset(info, 0, sizeof(*info));
  void* sc = nullptr;
  FRICC_TRY(compress_launch(x, n, m, method, u, nullptr, out_idx, out_val, nullptr, s, &sc));
  CompScalars h;
  FRICC_HIP(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s));
  FRICC_HIP(hipStreamSynchronize(s));
  if (dup) *dup = h.dup != 0;
  return comp_scalars_status(&h, info);
}

// ---------------------------------------------------------------- allocation-free entry
// fricc_compress without any allocation, for callers that cannot allocate
// (graph capture, tight serving loops): the whole scratch of the pipeline
// lives in this file-scope device array, of fixed size for vectors of up to
// NOALLOC_MAX_N elements.  Each GPU the library runs on holds its own copy of
// the array until the process ends; its address on the current device is
// looked up once (hipGetSymbolAddress) and kept per device -- in host code the
// array's name is not a device address.  Calls on one GPU share the array, so
// calls that overlap on one GPU race; every call initialises what it reads
// (scalars and histograms by kernels), so nothing carries over between calls.
constexpr int64_t NOALLOC_MAX_N = (int64_t)1 << 22;
// comp_layout(NOALLOC_MAX_N).total is 25.25 bytes per element plus a few
// hundred KB of small parts; the array has 26 per element plus 1 MB
constexpr size_t NOALLOC_BYTES = (size_t)26 * NOALLOC_MAX_N + ((size_t)1 << 20);
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAXDEV = 16;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAXDEV] = {nullptr};

// Device address of `scratch` on the current device (looked up on the first
// call on that device, with that device current).
static int scratch_address(void** out) {
  int dev = 0;
  FRICC_HIP(hipGetDevice(&dev));
  FRICC_CHECK(dev >= 0 && dev < NOALLOC_MAXDEV, "compress_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    FRICC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FRICC_CHECK(p != nullptr, "compress_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return OK;
}

int64_t compress_noalloc_max_n() { return NOALLOC_MAX_N; }

int compress_noalloc(const double* x, int64_t n, int64_t m, int method, double u, uint32_t* out_idx,
                     double* out_val, CompressInfo* info, hipStream_t s) {
  memset(info, 0, sizeof(*info));
  FRICC_TRY(check_compress_args(n, m, method, u, nullptr));
  FRICC_CHECK(n <= NOALLOC_MAX_N,
              "compress_noalloc: vector longer than the fixed scratch allows (at most 4194304 elements; "
              "use fricc_compress)");
  const CompLayout L = comp_layout(n);
  FRICC_CHECK(L.total <= NOALLOC_BYTES, "compress_noalloc: scratch layout exceeds the fixed array");
  void* base = nullptr;
  FRICC_TRY(scratch_address(&base));
  const CompScratch w = comp_scratch_at((uint8_t*)base, L);
  FRICC_TRY(compress_pipeline(x, n, m, method, u, nullptr, out_idx, out_val, nullptr, s, w));
  CompScalars h;
  FRICC_HIP(hipMemcpyAsync(&h, w.sc, sizeof(h), hipMemcpyDeviceToHost, s));
  FRICC_HIP(hipStreamSynchronize(s));
  return comp_scalars_status(&h, info);
}

// Systematic sampling of n_sample indices from given probabilities p[n]
// (each in [0,1]) with the canonical interval ends.
int sample_systematic(const double* p, int64_t n, int64_t n_sample, double u, uint32_t* out_idx,
                      int64_t* n_notfound, hipStream_t s) {
  FRICC_CHECK(n > 0 && n < (int64_t)1 << 31, "sample_systematic: length must be in [1, 2^31)");
  FRICC_CHECK(n_sample >= 0, "sample_systematic: negative sample count");
  FRICC_CHECK(u >= 0.0 && u <= 1.0, "sample_systematic: uniform must lie in [0, 1]");
  if (n_notfound) *n_notfound = 0;
  if (n_sample == 0) return OK;
  void *p_sc, *p_ts, *p_tot, *p_ths;
  const int64_t nt = ceil_div(n, TILE);
  FRICC_TRY(ws_get(WS_SC, sizeof(CompScalars), &p_sc));
  FRICC_TRY(ws_get(WS_TS, (size_t)(nt + 1) * sizeof(u128), &p_ts));
  FRICC_TRY(ws_get(WS_TOT, 64, &p_tot));
  FRICC_TRY(ws_get(WS_THS, (size_t)nt * NT * sizeof(u128), &p_ths));
  CompScalars* sc = (CompScalars*)p_sc;
  u128* ts = (u128*)p_ts;
  k_init_scalars<<<1, NT, 0, s>>>(sc, 0, n_sample, u, nullptr, nullptr, 0);
  k_samp_tilesums<<<nt, NT, 0, s>>>(nullptr, p, n, sc, ts, (u128*)p_ths);
  k_scan_u128_sums<<<1, 1024, 0, s>>>(ts, nt, (u128*)p_tot);
  EmitScratch es;
  FRICC_TRY(emit_scratch_ws(n, &es));
  samp_emit(nullptr, p, n, sc, ts, (const u128*)p_ths, out_idx, nullptr, es, nt, s);
  k_samp_final<<<1, 1, 0, s>>>(sc, (u128*)p_tot);
  k_samp_notfound<<<ceil_div(n_sample, NT), NT, 0, s>>>(p, sc, out_idx, nullptr);
  FRICC_HIP(hipGetLastError());
  CompScalars h;
  FRICC_HIP(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s));
  FRICC_HIP(hipStreamSynchronize(s));
  if (n_notfound) *n_notfound = (int64_t)h.n_notfound;
  if (h.pflag) {
    set_error("sample_systematic: a probability exceeds 1");
    return EPROB;
  }
  return OK;
}

// Pivotal sampling of n_sample indices from given probabilities p[n] (each
// in [0,1]); step j draws U[j] (device, n doubles; U[0] unused) or, with U
// null, the SplitMix64 stream keyed by the bits of u.
int sample_pivotal(const double* p, int64_t n, int64_t n_sample, const double* U, double u, uint32_t* out_idx,
                   hipStream_t s) {
  FRICC_CHECK(n > 0 && n < (int64_t)1 << 31, "sample_pivotal: length must be in [1, 2^31)");
  FRICC_CHECK(n_sample >= 0, "sample_pivotal: negative sample count");
  FRICC_CHECK(U || (u >= 0.0 && u <= 1.0), "sample_pivotal: uniform must lie in [0, 1]");
  void *p_sc, *p_ts, *p_tot, *p_ths;
  const int64_t nt = ceil_div(n, TILE);
  FRICC_TRY(ws_get(WS_SC, sizeof(CompScalars), &p_sc));
  FRICC_TRY(ws_get(WS_TS, (size_t)(nt + 1) * sizeof(u128), &p_ts));
  FRICC_TRY(ws_get(WS_TOT, 64, &p_tot));
  FRICC_TRY(ws_get(WS_THS, (size_t)nt * NT * sizeof(u128), &p_ths));
  uint32_t *tmax, *thmax;
  uint64_t* pst;
  FRICC_TRY(piv_scratch_ws(n, &tmax, &thmax, &pst));
  CompScalars* sc = (CompScalars*)p_sc;
  u128* ts = (u128*)p_ts;
  k_init_scalars<<<1, NT, 0, s>>>(sc, 0, n_sample, U ? 0.0 : u, nullptr, nullptr, 0);
  k_samp_tilesums<<<nt, NT, 0, s>>>(nullptr, p, n, sc, ts, (u128*)p_ths);
  k_scan_u128_sums<<<1, 1024, 0, s>>>(ts, nt, (u128*)p_tot);
  FRICC_TRY(pivotal_passes(nullptr, p, U, n, n_sample, sc, ts, (const u128*)p_ths, (const u128*)p_tot, tmax, thmax,
                           pst, out_idx, nullptr, s));
  CompScalars h;
  FRICC_HIP(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s));
  FRICC_HIP(hipStreamSynchronize(s));
  if (h.pflag) {
    set_error("sample_pivotal: a probability lies outside [0, 1]");
    return EPROB;
  }
  if (h.pivbad) {
    set_error("sample_pivotal: the chain ended with a sample count other than n_sample (reference exits, "
              "src/fri_utils.cpp:100-123)");
    return ESAMPLE;
  }
  return OK;
}

// parallel_sample of n_sample indices from q[n] with T chunks.
int parallel_sample(const double* q, int64_t n, int64_t n_sample, int64_t T, const double* U, double u,
                    uint32_t* out_idx, hipStream_t s) {
  FRICC_CHECK(n > 0 && n < (int64_t)1 << 31, "parallel_sample: length must be in [1, 2^31)");
  FRICC_CHECK(n_sample >= 0, "parallel_sample: negative sample count");
  FRICC_CHECK(T >= 1 && T <= n && T <= 65536, "parallel_sample: chunk count must lie in [1, min(n, 65536)]");
  FRICC_CHECK(U || (u >= 0.0 && u <= 1.0), "parallel_sample: uniform must lie in [0, 1]");
  const int64_t step = n / T;
  void *p_sc, *p_qj, *p_plan, *p_qa;
  FRICC_TRY(ws_get(WS_SC, sizeof(CompScalars), &p_sc));
  FRICC_TRY(ws_get(WS_TS, (size_t)T * sizeof(u128), &p_qj));
  FRICC_TRY(ws_get(WS_PST, (size_t)T * sizeof(PsPlan), &p_plan));
  FRICC_TRY(ws_get(WS_K0, (size_t)n * 8, &p_qa));
  CompScalars* sc = (CompScalars*)p_sc;
  const u128* QJ = (const u128*)p_qj;
  PsPlan* plan = (PsPlan*)p_plan;
  k_init_scalars<<<1, NT, 0, s>>>(sc, 0, n_sample, U ? 0.0 : u, nullptr, nullptr, 0);
  k_ps_chunksum<<<T, NT, 0, s>>>(q, n, T, step, (u128*)p_qj, sc);
  k_ps_plan<<<1, 64, 0, s>>>(QJ, T, step, U, plan, sc);
  k_ps_move<<<T, NT, 0, s>>>(q, n, T, step, QJ, plan, (double*)p_qa, sc);
  if (n_sample > 0) k_piv_clear<<<ceil_div(n_sample, NT), NT, 0, s>>>(sc, n_sample, out_idx, nullptr);
  if (U) {
    k_ps_chain<true><<<T, NT, 0, s>>>(q, (const double*)p_qa, U, n, T, step, n_sample, plan, sc, out_idx);
  } else {
    k_ps_chain<false><<<T, NT, 0, s>>>(q, (const double*)p_qa, U, n, T, step, n_sample, plan, sc, out_idx);
  }
  FRICC_HIP(hipGetLastError());
  CompScalars h;
  FRICC_HIP(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s));
  FRICC_HIP(hipStreamSynchronize(s));
  if (h.pflag) {
    set_error("parallel_sample: a probability lies outside [0, 1]");
    return EPROB;
  }
  if (h.pivbad) {
    set_error("parallel_sample: a chain ended with a sample count other than asked (reference exits, "
              "src/fri_utils.cpp:100-123)");
    return ESAMPLE;
  }
  return OK;
}

}  // namespace fricc
