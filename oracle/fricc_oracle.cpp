// CPU oracle for the FRI-CCSD hot path.  TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker (or as the timed CPU baseline).
// The product path (fri-cc_amd/) never links or calls it.
//
// It restates, by semantics, the reference's native hot path:
//   * FRI compression            src/fri_utils.cpp:11-36, 45-58, 129-159,
//                                164-171, 367-398, 400-437, 439-522
//   * SparseTensor4d             include/sparse_tensor.hpp:33-77, 102-143
//   * the 11 contraction kernels src/sparse_tensor.cpp:11-382
//
// Two compression modes:
//   CANONICAL (mode 0) - the definition the GPU implements bit-for-bit:
//     every sum that feeds a decision is an exact fixed-point sum, rounded
//     once to FP64 (round-to-nearest-even), so its value does not depend on
//     summation order or thread count.
//       - 1-norm R0: sum of FX(|x_i|) on the grid 2^G, G = max(eb_max,1)-1119
//         (eb_max = biased exponent of max|x|; elements are truncated at
//         2^-96 relative to the largest binade).
//       - D loop: R_k = round(Ftot - sum_{l<k} FX(y_sigma_l)); the test
//         (m-k)*y >= R_k - y && R_k > 1e-14 is evaluated in FP64 exactly as
//         src/fri_utils.cpp:415 does.
//       - systematic intervals: P_j = round(sum_{l<j} FXP(p_l) * 2^-96),
//         FXP(p) = floor(p * 2^96).
//     Ties in |x| are ordered by ascending flat index (what a stable merge
//     sort of (x_i, i) pairs built in index order yields, src/fri_utils.cpp:13-25).
//   FAITHFUL (mode 1) - the reference's own order of operations: sequential
//     one_norm, sequential subtraction in the D loop (src/fri_utils.cpp:411-421),
//     std::partial_sum intervals (src/fri_utils.cpp:142-143).
// Both modes then run the reference's forward scan over the intervals
// (src/fri_utils.cpp:147-157) literally, including the "not found -> index 0"
// behaviour of the zero-initialised S vector (src/fri_utils.cpp:132).
//
// Parity status of each function is recorded in the header comment of the
// function and pinned by tests/test_oracle_*.py.

#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <math.h>
#include <omp.h>
#include <immintrin.h>
#include <parallel/algorithm>

#include <vector>

#include <algorithm>
#include <numeric>
#include <vector>
#include <string>

typedef unsigned __int128 u128;

// (pq|rs) = fma chain over Q, tiled like a GEMM: rows (p,q) with their L
// values packed [row][Q], columns (r,s) in chunks packed [Q][col]; each output
// element is still acc = 0, acc = fma(a_Q, b_Q, acc) for Q = 0..naux-1 in
// order (vector lanes hold different elements), so every value is bitwise
// that of the plain loop.
__attribute__((target("avx2,fma"))) static void eri_block_packed(const double* L, int64_t naux, int64_t nmo,
                                                                  int64_t p0, int64_t p1, int64_t q0, int64_t q1,
                                                                  int64_t r0, int64_t r1, int64_t s0, int64_t s1,
                                                                  double* out) {
  const int64_t nq = q1 - q0, nr = r1 - r0, ns = s1 - s0;
  const int64_t nrow = (p1 - p0) * nq, ncol = nr * ns, n2 = nmo * nmo;
  std::vector<double> At((size_t)nrow * naux);
#pragma omp parallel for schedule(static)
  for (int64_t row = 0; row < nrow; row++) {
    const int64_t off = (row / nq + p0) * nmo + (row % nq + q0);
    for (int64_t Q = 0; Q < naux; Q++) At[(size_t)row * naux + Q] = L[Q * n2 + off];
  }
  constexpr int64_t CB = 240;  // columns per chunk: 20 groups of 12
  const int64_t nchunk = (ncol + CB - 1) / CB;
#pragma omp parallel
  {
    std::vector<double> Bp((size_t)naux * CB);
#pragma omp for schedule(dynamic, 1)
    for (int64_t ch = 0; ch < nchunk; ch++) {
      const int64_t c0 = ch * CB, nc = std::min(CB, ncol - c0);
      for (int64_t Q = 0; Q < naux; Q++)
        for (int64_t c = 0; c < CB; c++) {
          const int64_t col = c0 + c;
          Bp[(size_t)Q * CB + c] =
              c < nc ? L[Q * n2 + (col / ns + r0) * nmo + (col % ns + s0)] : 0.0;
        }
      int64_t row = 0;
      for (; row + 4 <= nrow; row += 4) {
        const double* a0 = &At[(size_t)row * naux];
        for (int64_t j = 0; j < nc; j += 12) {
          __m256d acc[4][3];
          for (int i = 0; i < 4; i++)
            for (int k = 0; k < 3; k++) acc[i][k] = _mm256_setzero_pd();
          for (int64_t Q = 0; Q < naux; Q++) {
            const double* b = &Bp[(size_t)Q * CB + j];
            const __m256d b0 = _mm256_loadu_pd(b), b1 = _mm256_loadu_pd(b + 4), b2 = _mm256_loadu_pd(b + 8);
            for (int i = 0; i < 4; i++) {
              const __m256d a = _mm256_broadcast_sd(a0 + (size_t)i * naux + Q);
              acc[i][0] = _mm256_fmadd_pd(a, b0, acc[i][0]);
              acc[i][1] = _mm256_fmadd_pd(a, b1, acc[i][1]);
              acc[i][2] = _mm256_fmadd_pd(a, b2, acc[i][2]);
            }
          }
          for (int i = 0; i < 4; i++) {
            alignas(32) double t[12];
            for (int k = 0; k < 3; k++) _mm256_store_pd(t + 4 * k, acc[i][k]);
            const int64_t w = std::min<int64_t>(12, nc - j);
            for (int64_t c = 0; c < w; c++) out[(size_t)(row + i) * ncol + c0 + j + c] = t[c];
          }
        }
      }
      for (; row < nrow; row++) {
        const double* a = &At[(size_t)row * naux];
        for (int64_t c = 0; c < nc; c++) {
          double acc = 0.0;
          for (int64_t Q = 0; Q < naux; Q++) acc = __builtin_fma(a[Q], Bp[(size_t)Q * CB + c], acc);
          out[(size_t)row * ncol + c0 + c] = acc;
        }
      }
    }
  }
}

extern "C" {

struct orc_info {
  int64_t n_D;            // |D|, exactly kept entries
  int64_t n_sampled;      // n' = m - |D|
  int64_t n_notfound;     // samples that fell outside every interval (index 0)
  double norm0;           // 1-norm before the D loop
  double remaining_norm;  // R after the D loop (denominator of p)
  double p_total;         // last interval end P_N
  int32_t status;         // 0 ok, 1 p>1 (reference exits), 2 bad args, 3 non-finite
  int32_t pad;
};

// ---------------------------------------------------------------------------
// Fixed-point helpers (canonical mode).  Pinned by tests/test_oracle_fri.py
// against Python's exact rational arithmetic (math.fsum / fractions).

static inline uint64_t dbits(double v) { uint64_t b; memcpy(&b, &v, 8); return b; }

// FX(y) on the grid 2^G with G = max(eb_max,1) - 1119: exact for every value
// within 2^44 of the top binade, truncated below.
static inline u128 fx_of(uint64_t ybits, uint32_t eb_max) {
  uint32_t eb = (uint32_t)((ybits >> 52) & 0x7ff);
  uint64_t M = (ybits & ((1ull << 52) - 1)) | (eb ? (1ull << 52) : 0ull);
  int s = (int)std::max(eb, 1u) - (int)std::max(eb_max, 1u) + 44;
  if (eb == 0x7ffu) return 0;
  if (s > 74) s = 74;
  if (s >= 0) return ((u128)M) << s;
  int r = -s;
  return r >= 64 ? (u128)0 : (u128)(M >> r);
}
static inline int fx_grid(uint32_t eb_max) { return (int)std::max(eb_max, 1u) - 1119; }

// Round-to-nearest-even conversion of v * 2^G to double.
static inline double u128_to_double(u128 v, int G) {
  if (v == 0) return 0.0;
  uint64_t hi = (uint64_t)(v >> 64), lo = (uint64_t)v;
  int nb = hi ? 128 - __builtin_clzll(hi) : 64 - __builtin_clzll(lo);
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
      if (mant == (1ull << 53)) { mant >>= 1; sh++; }
    }
    e = sh;
  }
  return ldexp((double)mant, e + G);
}

double orc_norm_canonical(const double* x, int64_t n) {
  uint32_t ebm = 0;
  for (int64_t i = 0; i < n; i++) ebm = std::max(ebm, (uint32_t)((dbits(x[i]) >> 52) & 0x7ff));
  u128 F = 0;
  for (int64_t i = 0; i < n; i++) F += fx_of(dbits(x[i]) & 0x7fffffffffffffffull, ebm);
  return u128_to_double(F, fx_grid(ebm));
}

// P[0..n] canonical interval ends for probabilities p (all in [0,1]).
void orc_prefix_canonical(const double* p, int64_t n, double* P) {
  u128 F = 0;
  P[0] = 0.0;
  for (int64_t j = 0; j < n; j++) {
    F += fx_of(dbits(p[j]), 1023u);
    P[j + 1] = u128_to_double(F, -96);
  }
}

// ---------------------------------------------------------------------------
// sigma: indices sorted by |x| descending, ties by ascending index.
// full_sort: sort all n (the reference's argsort merge-sorts the whole vector
// in parallel and keeps the first m, src/fri_utils.cpp:11-36; the FAITHFUL
// mode does the same work, in libstdc++'s parallel sort), else a partial sort.
static void sigma_sort(const double* x, int64_t n, std::vector<int64_t>& sig, int64_t keep, bool full_sort = false) {
  sig.resize(n);
  std::iota(sig.begin(), sig.end(), (int64_t)0);
  auto cmp = [x](int64_t a, int64_t b) {
    double fa = fabs(x[a]), fb = fabs(x[b]);
    if (fa != fb) return fa > fb;
    return a < b;
  };
  if (full_sort) {
    __gnu_parallel::sort(sig.begin(), sig.end(), cmp);
    if (keep < n) sig.resize(keep);
  } else if (keep < n) {
    std::partial_sort(sig.begin(), sig.begin() + keep, sig.end(), cmp);
    sig.resize(keep);
  } else {
    std::sort(sig.begin(), sig.end(), cmp);
  }
}

// Reference forward scan of sample_systematic (src/fri_utils.cpp:129-159)
// over the given interval ends P[0..n] (uk_i = i*1.0 + u, lin_space_add_const
// src/fri_utils.cpp:45-58 with step (n'-0)/n' == 1.0).
static int64_t systematic_scan(int64_t n_sample, const double* P, int64_t n, double u,
                               std::vector<int64_t>& S) {
  S.assign(n_sample, 0);
  int64_t last = 0, notfound = 0;
  for (int64_t i = 0; i < n_sample; i++) {
    double uk = (double)i * 1.0 + u;
    bool found = false;
    for (int64_t j = last; j < n; j++) {
      if (uk >= P[j] && uk < P[j + 1]) { S[i] = j; last = j; found = true; break; }
    }
    if (!found) notfound++;
  }
  return notfound;
}

// ---------------------------------------------------------------------------
// Pivotal sampling, sample_pivotal (src/fri_utils.cpp:63-127).
//
// Step j (j = 1 .. n-1) of the chain draws one uniform: U[j] when an array is
// given (U[0] is never read), else the j-th output of SplitMix64 seeded with
// `key` (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): z = key + (j+1) * 0x9e3779b97f4a7c15, the mix13
// finalizer, the top 53 bits times 2^-53.  The reference draws rand()/RAND_MAX
// at every step (:73); passing that stream as U reproduces its draws.
static inline uint64_t splitmix_at(uint64_t key, int64_t j) {
  uint64_t z = key + (uint64_t)(j + 1) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static inline double piv_uniform(const double* U, uint64_t key, int64_t j) {
  return U ? U[j] : (double)(splitmix_at(key, j) >> 11) * 0x1.0p-53;
}

void orc_splitmix_uniforms(uint64_t key, int64_t n, double* out) {
  for (int64_t j = 0; j < n; j++) out[j] = piv_uniform(nullptr, key, j);
}

// The chain: a candidate i (initially 0) and the running mass a.  At a step
// that does not complete a unit of mass (a + p_j < 1) the candidate becomes j
// with probability p_j / a; at one that does, a drops by 1 and either the
// candidate is emitted and replaced by j (probability (1 - p_j) / (1 - a)) or
// j is emitted.  The loop stops once n_sample indices are out; at the last
// step a mass within 1e-5 of 1 emits the candidate.  Every other ending, and
// an acceptance probability outside [0, 1], is an exit of the reference
// (:85-98, :103-123): status 5.
//   FAITHFUL  (mode 1): a is the reference's FP64 running sum (:72-84).
//   CANONICAL (mode 0): the mass comes from the exact prefix
//     Q_{j+1} = sum_{l<=j} FXP(p_l) (units of 2^-96): the units completed by
//     step j are c_j = min(j, floor(Q_{j+1})) and a_j = round(Q_{j+1} - c_j);
//     the reference's sequence in exact arithmetic (the min covers p_0 = 1,
//     which the reference carries as a = 1 into step 1).
// p outside [0, 1] (or NaN): status 1.  S receives min(count, n_sample)
// indices, *cnt the number the chain emitted.
int orc_pivotal(const double* p, int64_t n, int64_t ns, const double* U, uint64_t key, int mode, int64_t* S,
                int64_t* cnt) {
  *cnt = 0;
  if (n <= 0 || ns < 0) return 2;
  for (int64_t j = 0; j < n; j++)
    if (!(p[j] >= 0.0 && p[j] <= 1.0)) return 1;
  int64_t i = 0, k = 0, c = 0;
  double a = p[0];
  u128 Q = fx_of(dbits(p[0]), 1023u);
  auto push = [&](int64_t v) {
    if (k < ns) S[k] = v;
    k++;
  };
  for (int64_t j = 1; j < n; j++) {
    const double rn = piv_uniform(U, key, j);
    bool cross;
    if (mode == 1) {
      a += p[j];
      cross = !(a < 1.0);
      if (cross) a -= 1.0;
    } else {
      Q += fx_of(dbits(p[j]), 1023u);
      const int64_t cn = std::min<int64_t>(j, (int64_t)(Q >> 96));
      cross = cn > c;
      c = cn;
      a = u128_to_double(Q - ((u128)c << 96), -96);
    }
    if (!cross) {
      if (rn < p[j] / a) i = j;
    } else {
      const double pa = (1.0 - p[j]) / (1.0 - a);
      if (pa > 1.0 || pa < 0.0) { *cnt = k; return 5; }
      if (rn < pa) {
        push(i);
        i = j;
      } else {
        push(j);
      }
    }
    if (k == ns) break;
    if (j == n - 1) {
      if (fabs(a - 1.0) < 1e-5) push(i);
      else { *cnt = k; return 5; }
    }
  }
  *cnt = k;
  return k == ns ? 0 : 5;
}

// ---------------------------------------------------------------------------
// Thread-partitioned pivotal sampling, parallel_sample (src/fri_utils.cpp:176-362),
// with T chunks in place of the reference's T OpenMP threads: chunk j is
// [j*step, (j+1)*step) (the last one takes the remainder), step = n / T.
//   a_j = n' * |q_j| / |q| (FP64, that order), t_j = a_j - floor(a_j),
//   g_j = floor(a_j); c = n' - sum g_j more samples go to the chunks a pivotal
//   chain over t picks (g_j += 1); each chunk's probabilities are then moved
//   in index order toward a sum of g_j -- up: y = min(1, q/t_j), down:
//   y = max(0, (q - t_j)/(1 - t_j)) -- until the running sum crosses g_j,
//   where the crossing element takes what makes the sum g_j exactly; the
//   chunk then runs its own pivotal chain for g_j samples, written at the
//   chunk's offset (prefix of g) with indices shifted by j*step.
// Draws: the chunk chain's step j takes the uniform of position j*step, the
// chain of chunk j its step i the uniform of position j*step + i (positions
// never shared; piv_uniform).
// CANONICAL (mode 0): |q| and |q_j| are the exact sums of FXP(q) rounded
// once, the running sum of the move is exact (FXP units), the crossing
// element takes round(FXP(y) + g - S), and every chain is the canonical one.
// FAITHFUL (mode 1): the reference's FP64 sums (sequential) and chains.
// Returns 0, 1 (q outside [0, 1]), 2 (T outside [1, n]) or 5 (a reference
// exit in one of the chains, or fewer than 0 extra samples).
int orc_parallel_sample(const double* q, int64_t n, int64_t ns, int64_t T, const double* U, uint64_t key, int mode,
                        int64_t* S) {
  if (n <= 0 || ns < 0 || T < 1 || T > n) return 2;
  for (int64_t i = 0; i < n; i++)
    if (!(q[i] >= 0.0 && q[i] <= 1.0)) return 1;
  const int64_t step = n / T, rem = n % T;
  std::vector<int64_t> lo(T), hi(T), g(T);
  std::vector<u128> QJ(T, 0);
  std::vector<double> qjn(T), a(T), t(T);
  u128 Qt = 0;
  double qn = 0.0;
  for (int64_t j = 0; j < T; j++) {
    lo[j] = j * step;
    hi[j] = (j + 1) * step + (j == T - 1 ? rem : 0);
    double sj = 0.0;
    for (int64_t i = lo[j]; i < hi[j]; i++) {
      QJ[j] += fx_of(dbits(q[i]), 1023u);
      sj += fabs(q[i]);
    }
    Qt += QJ[j];
    qjn[j] = mode == 1 ? sj : u128_to_double(QJ[j], -96);
  }
  if (mode == 1) {
    for (int64_t i = 0; i < n; i++) qn += fabs(q[i]);
  } else {
    qn = u128_to_double(Qt, -96);
  }
  if (qn == 0.0) return ns == 0 ? 0 : 5;
  int64_t c = ns;
  for (int64_t j = 0; j < T; j++) {
    a[j] = (double)ns * qjn[j] / qn;
    t[j] = a[j] - floor(a[j]);
    g[j] = (int64_t)floor(a[j]);
    c -= g[j];
  }
  if (c < 0) return 5;
  {
    std::vector<double> Uc(T);
    for (int64_t j = 0; j < T; j++) Uc[j] = piv_uniform(U, key, j * step);
    std::vector<int64_t> sp(std::max<int64_t>(c, 1));
    int64_t cnt = 0;
    const int st = orc_pivotal(t.data(), T, c, Uc.data(), 0, mode, sp.data(), &cnt);
    if (st != 0) return st;
    for (int64_t k = 0; k < c; k++) g[sp[k]] += 1;
  }
  int64_t lb = 0;
  for (int64_t j = 0; j < T; j++) {
    const int64_t len = hi[j] - lo[j];
    std::vector<double> qa(q + lo[j], q + hi[j]);
    const bool up = (double)g[j] > a[j], down = (double)g[j] < a[j];
    if (up || down) {
      const double tj = t[j];
      u128 Sx = QJ[j];
      const u128 G = (u128)g[j] << 96;
      double sj = qjn[j];
      for (int64_t i = 0; i < len; i++) {
        double y;
        if (up) {
          const double r = qa[i] / tj;
          y = r < 1.0 ? r : 1.0;
        } else {
          const double r = (qa[i] - tj) / (1.0 - tj);
          y = 0.0 < r ? r : 0.0;
        }
        bool cross;
        if (mode == 1) {
          sj += y - qa[i];
          cross = up ? sj >= (double)g[j] : sj <= (double)g[j];
          qa[i] = y;
          if (cross) qa[i] = y + (double)g[j] - sj;
        } else {
          const u128 Fy = fx_of(dbits(y), 1023u);
          Sx = Sx - fx_of(dbits(qa[i]), 1023u) + Fy;
          cross = up ? Sx >= G : Sx <= G;
          qa[i] = y;
          // Fy + G - S, negative when a chunk already held more than g (q
          // not summing to n'): the chain then rejects it as outside [0, 1]
          if (cross) qa[i] = Fy + G >= Sx ? u128_to_double(Fy + G - Sx, -96) : -u128_to_double(Sx - Fy - G, -96);
        }
        if (cross) break;
      }
    }
    std::vector<double> Uj(len);
    for (int64_t i = 0; i < len; i++) Uj[i] = piv_uniform(U, key, lo[j] + i);
    std::vector<int64_t> Sj(std::max<int64_t>(g[j], 1));
    int64_t cnt = 0;
    const int st = orc_pivotal(qa.data(), len, g[j], Uj.data(), 0, mode, Sj.data(), &cnt);
    if (st != 0) return st;
    for (int64_t k = 0; k < g[j]; k++) S[lb + k] = Sj[k] + j * step;
    lb += g[j];
  }
  return 0;
}

// FRI / largest compression.  method 0 = "largest" (include/sparse_tensor.hpp:46-56),
// method 1 = "fri" + "systematic" (include/sparse_tensor.hpp:58-67), method 2 =
// "fri" + "pivotal" (src/fri_utils.cpp:486-489), its step draws SplitMix64
// keyed by the bits of u (piv_uniform).
int orc_compress(const double* x, int64_t n, int64_t m, int method, int mode, double u,
                 int64_t* out_idx, double* out_val, orc_info* info) {
  memset(info, 0, sizeof(*info));
  if (n <= 0 || m <= 0 || m > n) { info->status = 2; return 2; }
  for (int64_t i = 0; i < n; i++)
    if (!std::isfinite(x[i])) { info->status = 3; return 3; }
  std::vector<int64_t> sig;
  if (method == 0) {
    sigma_sort(x, n, sig, m, mode == 1);
    for (int64_t k = 0; k < m; k++) { out_idx[k] = sig[k]; out_val[k] = x[sig[k]]; }
    info->n_D = m;
    return 0;
  }
  // get_d_largest (src/fri_utils.cpp:400-437)
  sigma_sort(x, n, sig, m, mode == 1);
  int64_t K = m;
  double R = 0.0;
  if (mode == 0) {
    uint32_t ebm = 0;
    for (int64_t i = 0; i < n; i++) ebm = std::max(ebm, (uint32_t)((dbits(x[i]) >> 52) & 0x7ff));
    const int G = fx_grid(ebm);
    u128 Ftot = 0;
    for (int64_t i = 0; i < n; i++) Ftot += fx_of(dbits(x[i]) & 0x7fffffffffffffffull, ebm);
    info->norm0 = u128_to_double(Ftot, G);
    u128 Fpre = 0;
    for (int64_t k = 0; k < m; k++) {
      double Rk = u128_to_double(Ftot - Fpre, G);
      double y = fabs(x[sig[k]]);
      if ((double)(m - k) * y >= Rk - y && Rk > 1e-14) {
        Fpre += fx_of(dbits(y), ebm);
      } else { K = k; break; }
    }
    R = u128_to_double(Ftot - Fpre, G);
  } else {
    double r = 0.0;
    for (int64_t i = 0; i < n; i++) r += fabs(x[i]);
    info->norm0 = r;
    for (int64_t k = 0; k < m; k++) {
      double y = fabs(x[sig[k]]);
      if ((double)(m - k) * y >= r - y && r > 1e-14) { r -= y; }
      else { K = k; break; }
    }
    R = r;
  }
  info->n_D = K;
  info->remaining_norm = R;
  const int64_t ns = m - K;
  info->n_sampled = ns;
  // make_probability_vector (src/fri_utils.cpp:367-398)
  std::vector<double> p(n);
  const double nsd = (double)ns;
  for (int64_t i = 0; i < n; i++) p[i] = fabs(x[i]) / R * nsd;
  for (int64_t k = 0; k < K; k++) p[sig[k]] = 0.0;
  // p > 1 makes the reference exit (src/fri_utils.cpp:383-395); a NaN p
  // (remaining norm 0 when m exceeds the nonzeros, see the guard at :415)
  // makes it emit index 0 with a non-finite value.  Both are reported.  With
  // n' = 0 (D took all m entries) nothing is sampled and p is never used: the
  // reference's check passes (NaN > 1 is false) and it returns D alone.
  std::vector<int64_t> S;
  if (ns > 0) {
    for (int64_t i = 0; i < n; i++)
      if (!(p[i] <= 1.0)) { info->status = 1; return 1; }
    std::vector<double> P(n + 1);
    if (mode == 0) {
      orc_prefix_canonical(p.data(), n, P.data());
    } else {
      P[0] = 0.0;
      std::partial_sum(p.begin(), p.end(), P.begin() + 1);
    }
    info->p_total = P[n];
    if (method == 2) {
      S.assign(ns, 0);
      int64_t cnt = 0;
      if (orc_pivotal(p.data(), n, ns, nullptr, dbits(u), mode, S.data(), &cnt) != 0) {
        info->status = 5;
        return 5;
      }
    } else {
      info->n_notfound = systematic_scan(ns, P.data(), n, u, S);
    }
  } else {
    info->p_total = 0.0;
    info->n_notfound = 0;
  }
  // wrap-up (src/fri_utils.cpp:494-504)
  for (int64_t k = 0; k < K; k++) { out_idx[k] = sig[k]; out_val[k] = x[sig[k]]; }
  for (int64_t i = 0; i < ns; i++) {
    int64_t j = S[i];
    out_idx[K + i] = j;
    out_val[K + i] = x[j] / p[j];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// The 11 contraction kernels, in the loop structure of src/sparse_tensor.cpp
// (output-block parallel loops that rescan every sparse entry).  Used as the
// timed CPU baseline and as a restatement checked against np.einsum by
// tests/test_oracle_contract.py (the reference's own test identities,
// fricc/tests/test_sparse_tensors.py:48-269).
//
// idx4: nnz x 4 unpacked indices (SparseTensor4d::unpack_idx,
// include/sparse_tensor.hpp:108-125), val: nnz values.  Strides in elements.
// block_lo/block_hi restrict the outer collapsed loop to a sub-range of
// output blocks (used by the bounded CPU-baseline sample); pass 0/-1 for all.

#define W4(a, b, c, d) W[(a) * ws[0] + (b) * ws[1] + (c) * ws[2] + (d) * ws[3]]
#define O4(a, b, c, d) out[(a) * os[0] + (b) * os[1] + (c) * os[2] + (d) * os[3]]

int orc_contract_ref(const char* term, const double* W, const int64_t* wshape, const int64_t* ws,
                     const int64_t* idx4, const double* val, int64_t nnz, const int64_t* tdims,
                     double* out, const int64_t* os, int64_t block_lo, int64_t block_hi) {
  const int64_t no = tdims[0], nv = tdims[2];
  std::string t(term);
  auto run = [&](int64_t d0, int64_t d1, auto body) {
    int64_t nb = d0 * d1;
    int64_t lo = std::max<int64_t>(0, block_lo), hi = block_hi < 0 ? nb : std::min(nb, block_hi);
#pragma omp parallel for schedule(dynamic)
    for (int64_t blk = lo; blk < hi; blk++) body(blk / d1, blk % d1);
  };
  if (t == "0101") {  // src/sparse_tensor.cpp:12-37
    run(nv, nv, [&](int64_t a, int64_t b) {
      for (int64_t s = 0; s < nnz; s++) {
        const int64_t* id = idx4 + 4 * s;
        if (id[2] == a && id[3] == b) {
          int64_t k = id[0], l = id[1];
          for (int64_t i = 0; i < no; i++)
            for (int64_t j = 0; j < no; j++) O4(i, j, a, b) += W4(k, l, i, j) * val[s];
        }
      }
    });
  } else if (t == "2323") {  // :40-86
    const int64_t da = wshape[0], db = wshape[1];
    run(tdims[0], tdims[1], [&](int64_t i, int64_t j) {
      for (int64_t s = 0; s < nnz; s++) {
        const int64_t* id = idx4 + 4 * s;
        if (id[0] == i && id[1] == j) {
          int64_t c = id[2], d = id[3];
          for (int64_t a = 0; a < da; a++)
            for (int64_t b = 0; b < db; b++) O4(i, j, a, b) += W4(a, b, c, d) * val[s];
        }
      }
    });
  } else if (t == "1302") {  // :89-115
    run(nv, no, [&](int64_t b, int64_t j) {
      for (int64_t s = 0; s < nnz; s++) {
        const int64_t* id = idx4 + 4 * s;
        if (id[3] == b && id[1] == j) {
          int64_t k = id[0], c = id[2];
          for (int64_t i = 0; i < no; i++)
            for (int64_t a = 0; a < nv; a++) O4(i, j, a, b) += 2 * W4(a, k, i, c) * val[s];
        }
      }
    });
  } else if (t == "1202") {  // :118-144
    run(nv, no, [&](int64_t b, int64_t j) {
      for (int64_t s = 0; s < nnz; s++) {
        const int64_t* id = idx4 + 4 * s;
        if (id[3] == b && id[1] == j) {
          int64_t k = id[0], c = id[2];
          for (int64_t i = 0; i < no; i++)
            for (int64_t a = 0; a < nv; a++) O4(i, j, a, b) -= W4(a, k, c, i) * val[s];
        }
      }
    });
  } else if (t == "1303") {  // :147-173
    run(nv, no, [&](int64_t b, int64_t j) {
      for (int64_t s = 0; s < nnz; s++) {
        const int64_t* id = idx4 + 4 * s;
        if (id[2] == b && id[1] == j) {
          int64_t k = id[0], c = id[3];
          for (int64_t i = 0; i < no; i++)
            for (int64_t a = 0; a < nv; a++) O4(i, j, a, b) += W4(a, k, i, c) * val[s];
        }
      }
    });
  } else if (t == "1203") {  // :176-202
    run(nv, no, [&](int64_t a, int64_t j) {
      for (int64_t s = 0; s < nnz; s++) {
        const int64_t* id = idx4 + 4 * s;
        if (id[2] == a && id[1] == j) {
          int64_t k = id[0], c = id[3];
          for (int64_t i = 0; i < no; i++)
            for (int64_t b = 0; b < nv; b++) O4(i, j, a, b) += W4(b, k, c, i) * val[s];
        }
      }
    });
  } else if (t == "1323") {  // :205-230
    run(no, no, [&](int64_t i, int64_t j) {
      for (int64_t s = 0; s < nnz; s++) {
        const int64_t* id = idx4 + 4 * s;
        if (id[0] == i && id[1] == j) {
          int64_t c = id[2], d = id[3];
          for (int64_t k = 0; k < no; k++)
            for (int64_t l = 0; l < no; l++) O4(k, l, i, j) += W4(k, c, l, d) * val[s];
        }
      }
    });
  } else if (t == "0112") {  // :233-259
    run(nv, no, [&](int64_t a, int64_t i) {
      for (int64_t s = 0; s < nnz; s++) {
        const int64_t* id = idx4 + 4 * s;
        if (id[3] == a && id[0] == i) {
          int64_t l = id[1], d = id[2];
          for (int64_t k = 0; k < no; k++)
            for (int64_t c = 0; c < nv; c++) O4(a, k, i, c) -= 0.5 * W4(l, d, k, c) * val[s];
        }
      }
    });
  } else if (t == "0313") {  // :262-289
    run(nv, no, [&](int64_t a, int64_t i) {
      for (int64_t s = 0; s < nnz; s++) {
        const int64_t* id = idx4 + 4 * s;
        if (id[2] == a && id[0] == i) {
          int64_t l = id[1], d = id[3];
          for (int64_t k = 0; k < no; k++)
            for (int64_t c = 0; c < nv; c++) O4(a, k, i, c) -= 0.5 * W4(l, c, k, d) * val[s];
        }
      }
    });
  } else if (t == "0113") {  // :292-317
    run(nv, no, [&](int64_t a, int64_t i) {
      for (int64_t s = 0; s < nnz; s++) {
        const int64_t* id = idx4 + 4 * s;
        if (id[2] == a && id[0] == i) {
          int64_t l = id[1], d = id[3];
          for (int64_t k = 0; k < no; k++)
            for (int64_t c = 0; c < nv; c++) O4(a, k, i, c) += W4(l, d, k, c) * val[s];
        }
      }
    });
  } else if (t == "0312") {  // :320-346
    run(nv, no, [&](int64_t a, int64_t i) {
      for (int64_t s = 0; s < nnz; s++) {
        const int64_t* id = idx4 + 4 * s;
        if (id[3] == a && id[0] == i) {
          int64_t l = id[1], d = id[2];
          for (int64_t k = 0; k < no; k++)
            for (int64_t c = 0; c < nv; c++) O4(a, k, c, i) -= 0.5 * W4(l, c, k, d) * val[s];
        }
      }
    });
  } else {
    return 2;  // the reference exits (src/sparse_tensor.cpp:378-381)
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Two-electron integrals (pq|rs) = sum_Q L[Q,p,q] L[Q,r,s] as a sequential
// fused-multiply-add chain over Q = 0..naux-1 starting from 0.  Because
// fma(a,b,c) == fma(b,a,c) and L is bitwise symmetric in (p,q), every member
// of an 8-fold orbit gets bitwise the same value.  The GPU builder uses the
// same chain, so the integrals are bit-identical between the two.
// Writes out[(p-p0), (q-q0), (r-r0), (s-s0)] row-major.
// The plain loop: each element's fma chain over Q, L read with stride nmo^2.
void orc_eri_block_plain(const double* L, int64_t naux, int64_t nmo, int64_t p0, int64_t p1,
                         int64_t q0, int64_t q1, int64_t r0, int64_t r1, int64_t s0, int64_t s1,
                         double* out) {
  const int64_t np = p1 - p0, nq = q1 - q0, nr = r1 - r0, ns = s1 - s0;
  const int64_t n2 = nmo * nmo;
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t p = 0; p < np; p++)
    for (int64_t q = 0; q < nq; q++)
      for (int64_t r = 0; r < nr; r++)
        for (int64_t s = 0; s < ns; s++) {
          const double* a = L + (p + p0) * nmo + (q + q0);
          const double* b = L + (r + r0) * nmo + (s + s0);
          double acc = 0.0;
          for (int64_t Q = 0; Q < naux; Q++) acc = fma(a[Q * n2], b[Q * n2], acc);
          out[((p * nq + q) * nr + r) * ns + s] = acc;
        }
}

void orc_eri_block(const double* L, int64_t naux, int64_t nmo, int64_t p0, int64_t p1,
                   int64_t q0, int64_t q1, int64_t r0, int64_t r1, int64_t s0, int64_t s1,
                   double* out) {
  if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma"))
    eri_block_packed(L, naux, nmo, p0, p1, q0, q1, r0, r1, s0, s1, out);
  else
    orc_eri_block_plain(L, naux, nmo, p0, p1, q0, q1, r0, r1, s0, s1, out);
}

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
