// Dense contractions of two strided tensors on MFMA: the O(o^2 v^3) products
// of the amplitude update (PySCF's imd.cc_Foo / cc_Fvv / cc_Fov / Loo / Lvv,
// cc_Woooo / cc_Wvoov / cc_Wvovo and the direct t1/t2 terms,
// fricc/friccsd.py:144-207, 309-376) written as one batched GEMM each.
//
//   C[b][m][n] = alpha * sum_k A[b][m][k] B[b][k][n] + beta * C[b][m][n]
//
// Every index (b, m, n, k) stands for up to three letters, each with its own
// stride, so permuted views, broadcasts (stride 0) and products summed over
// several letters need no copies.  An einsum-style planner (dense_contract)
// groups the letters of a spec into b/m/n/k and merges adjacent letters whose
// strides compose.
//
// Tiling: 256 threads (4 waves as 2 x 2), a BM x BN output tile per block,
// K advancing 16 per step.  Each step's A (BK x BM) and B (BK x BN) tiles are
// loaded from global memory into registers while the MFMAs of the previous
// step run from LDS, then written to the other LDS buffer (one barrier per
// step).  Consecutive threads load consecutive elements along whichever of
// the tile's two indices has unit stride, so the loads coalesce.  Elements
// are converted to the compute type on load (FP64 operands of an FP32
// product need no converted copy).  Compute types: FP32 on
// v_mfma_f32_16x16x4f32, FP64 on v_mfma_f64_16x16x4f64; the accumulation of
// each output is a fixed sequence (no atomics), so results are reproducible.
// Small output grids split K; the partial tiles are added in split order by a
// second kernel.
#include "fricc_internal.h"
#include "mfma.h"
#include "dma.h"

#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

namespace fricc {

// An index i of a group of up to three letters: i = (i2 * L1 + i1) * L0 + i0,
// addressed as i2 * s2 + i1 * s1 + i0 * s0 (i < 2^31).
struct Lvl {
  int64_t L0, L1;
  int64_t s0, s1, s2;
  double r0, r1;  // 1 / L0, 1 / L1
  int nl;         // levels in use (0..3)
};

// q = i / L, r = i % L for 0 <= i < 2^31 via the FP64 reciprocal (the product
// is within one of the quotient; one correction step)
__device__ inline void qr(int64_t i, int64_t L, double rcp, int64_t& q, int64_t& r) {
  q = (int64_t)((double)i * rcp);
  r = i - q * L;
  if (r < 0) {
    q--;
    r += L;
  } else if (r >= L) {
    q++;
    r -= L;
  }
}

__device__ inline int64_t lvl_off(const Lvl& d, int64_t i) {
  if (d.nl <= 1) return i * d.s0;
  int64_t q, r0;
  qr(i, d.L0, d.r0, q, r0);
  if (d.nl == 2) return q * d.s1 + r0 * d.s0;
  int64_t q2, r1;
  qr(q, d.L1, d.r1, q2, r1);
  return q2 * d.s2 + r1 * d.s1 + r0 * d.s0;
}

struct DArgs {
  const void* A;
  const void* B;
  void* C;
  const void* Csrc;  // the beta term's source: C itself, or a tensor of C's layout
  void* partial;
  int a64, b64, c64;
  int a_kfast, b_kfast;  // the tile load walks k fastest (else m / n fastest)
  int c_vec;             // C rows hold runs of 4 consecutive, 16-byte aligned columns
  Lvl Ab, Am, Ak, Bb, Bk, Bn, Cb, Cm, Cn;
  int64_t M, N, K, batch;
  int64_t ntm, ntn;
  int nsplit;
  int64_t kchunk;  // K range per split, a multiple of DBK
  double alpha, beta;
};

constexpr int DBK = 16;

template <typename T>
__device__ inline T ld_as(const void* p, int f64, int64_t off) {
  return f64 ? (T)((const double*)p)[off] : (T)((const float*)p)[off];
}

// V consecutive elements at off (16-byte aligned when V == 4: the host
// checks that every stride but the unit one is a multiple of 4 and the base
// is aligned), converted to T
template <typename T, int V>
__device__ inline void ld_vec(const void* p, int f64, int64_t off, T (&r)[V]) {
  if constexpr (V == 4) {
    if (f64) {
      const double2 x = *(const double2*)((const double*)p + off);
      const double2 y = *(const double2*)((const double*)p + off + 2);
      r[0] = (T)x.x;
      r[1] = (T)x.y;
      r[2] = (T)y.x;
      r[3] = (T)y.y;
    } else {
      const float4 x = *(const float4*)((const float*)p + off);
      r[0] = (T)x.x;
      r[1] = (T)x.y;
      r[2] = (T)x.z;
      r[3] = (T)x.w;
    }
  } else {
    r[0] = ld_as<T>(p, f64, off);
  }
}

// VA / VB: elements per load along the unit-stride index of A / B (4 or 1);
// kfast flags pick which index that is (k, else m / n).
template <typename T, int BM, int BN, int VA, int VB>
__global__ __launch_bounds__(256) void k_dgemm(DArgs a) {
  constexpr int PA = BM + (sizeof(T) == 4 ? 16 : 0);  // LDS pitches: the two k rows one fragment read
  constexpr int PB = BN + (sizeof(T) == 4 ? 16 : 0);  // touches fall on different banks
  constexpr int NGA = BM * DBK / VA, NGB = BN * DBK / VB;  // load groups of a tile
  constexpr int GA = (NGA + 255) / 256, GB = (NGB + 255) / 256;
  constexpr int MB = BM / 32, NB = BN / 32;  // 16 x 16 blocks per wave
  typedef typename Acc<T>::type AT;
  constexpr int WN = BN / 2, EP = WN + 4;  // epilogue: a wave's 16-row slab, padded
  constexpr int SA_BYTES = 2 * DBK * PA * (int)sizeof(T), SB_BYTES = 2 * DBK * PB * (int)sizeof(T);
  static_assert(4 * 16 * EP * (int)sizeof(T) <= SA_BYTES + SB_BYTES, "epilogue slabs must fit the staging LDS");
  __shared__ __align__(16) unsigned char smem[SA_BYTES + SB_BYTES];
  T(*sA)[DBK][PA] = (T(*)[DBK][PA])smem;
  T(*sB)[DBK][PB] = (T(*)[DBK][PB])(smem + SA_BYTES);
  __shared__ int64_t cmoff[BM], cnoff[BN];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int64_t tm = (int64_t)blockIdx.x % a.ntm, tn = (int64_t)blockIdx.x / a.ntm;
  const int64_t b = (int64_t)blockIdx.y / a.nsplit;
  const int sp = (int)((int64_t)blockIdx.y % a.nsplit);
  const int64_t m0 = tm * BM, n0 = tn * BN;
  const int64_t kbeg = (int64_t)sp * a.kchunk, kend = min(a.K, kbeg + a.kchunk);
  const int64_t abase = lvl_off(a.Ab, b), bbase = lvl_off(a.Bb, b);
  for (int t = tid; t < BM; t += 256) cmoff[t] = (m0 + t < a.M) ? lvl_off(a.Cm, m0 + t) : 0;
  for (int t = tid; t < BN; t += 256) cnoff[t] = (n0 + t < a.N) ? lvl_off(a.Cn, n0 + t) : 0;

  // staging coordinates of this thread's load groups: (row index mm / nn,
  // first k) of V elements along the unit-stride index
  int amm[GA], akk[GA], bnn[GB], bkk[GB];
  int64_t amoff[GA], bnoff[GB];
  bool amok[GA], bnok[GB];
#pragma unroll
  for (int j = 0; j < GA; j++) {
    const int g = tid + 256 * j;
    const bool in = g < NGA;
    if (a.a_kfast) {
      amm[j] = g / (DBK / VA);
      akk[j] = (g % (DBK / VA)) * VA;
    } else {
      amm[j] = (g % (BM / VA)) * VA;
      akk[j] = g / (BM / VA);
    }
    const int64_t m = m0 + amm[j];
    amok[j] = in && m < a.M;
    amoff[j] = amok[j] ? abase + lvl_off(a.Am, m) : 0;
  }
#pragma unroll
  for (int j = 0; j < GB; j++) {
    const int g = tid + 256 * j;
    const bool in = g < NGB;
    if (a.b_kfast) {
      bnn[j] = g / (DBK / VB);
      bkk[j] = (g % (DBK / VB)) * VB;
    } else {
      bnn[j] = (g % (BN / VB)) * VB;
      bkk[j] = g / (BN / VB);
    }
    const int64_t n = n0 + bnn[j];
    bnok[j] = in && n < a.N;
    bnoff[j] = bnok[j] ? bbase + lvl_off(a.Bn, n) : 0;
  }
  T ra[GA][VA], rb[GB][VB];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int j = 0; j < GA; j++) {
      const int64_t k = k0 + akk[j];
      if (amok[j] && k < kend) {
        ld_vec<T, VA>(a.A, a.a64, amoff[j] + lvl_off(a.Ak, k), ra[j]);
      } else {
#pragma unroll
        for (int i = 0; i < VA; i++) ra[j][i] = (T)0;
      }
    }
#pragma unroll
    for (int j = 0; j < GB; j++) {
      const int64_t k = k0 + bkk[j];
      if (bnok[j] && k < kend) {
        ld_vec<T, VB>(a.B, a.b64, bnoff[j] + lvl_off(a.Bk, k), rb[j]);
      } else {
#pragma unroll
        for (int i = 0; i < VB; i++) rb[j][i] = (T)0;
      }
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int j = 0; j < GA; j++) {
      if (tid + 256 * j >= NGA) continue;
#pragma unroll
      for (int i = 0; i < VA; i++) {
        if (a.a_kfast)
          sA[buf][akk[j] + i][amm[j]] = ra[j][i];
        else
          sA[buf][akk[j]][amm[j] + i] = ra[j][i];
      }
    }
#pragma unroll
    for (int j = 0; j < GB; j++) {
      if (tid + 256 * j >= NGB) continue;
#pragma unroll
      for (int i = 0; i < VB; i++) {
        if (a.b_kfast)
          sB[buf][bkk[j] + i][bnn[j]] = rb[j][i];
        else
          sB[buf][bkk[j]][bnn[j] + i] = rb[j][i];
      }
    }
  };

  AT acc[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; i++)
#pragma unroll
    for (int j = 0; j < NB; j++) acc[i][j] = AT{0, 0, 0, 0};

  const int lr = lane & 15, lk = lane >> 4;
  const int mw = wm * (BM / 2), nw = wn * (BN / 2);
  if (kbeg < kend) {
    gload(kbeg);
    lstore(0);
    __syncthreads();
    int cur = 0;
    for (int64_t k0 = kbeg; k0 < kend; k0 += DBK) {
      const bool more = k0 + DBK < kend;
      if (more) gload(k0 + DBK);
#pragma unroll
      for (int ks = 0; ks < DBK / 4; ks++) {
        T af[MB], bf[NB];
#pragma unroll
        for (int i = 0; i < MB; i++) af[i] = sA[cur][4 * ks + lk][mw + 16 * i + lr];
#pragma unroll
        for (int j = 0; j < NB; j++) bf[j] = sB[cur][4 * ks + lk][nw + 16 * j + lr];
#pragma unroll
        for (int i = 0; i < MB; i++)
#pragma unroll
          for (int j = 0; j < NB; j++) acc[i][j] = Acc<T>::mfma(af[i], bf[j], acc[i][j]);
      }
      if (more) lstore(cur ^ 1);
      __syncthreads();
      cur ^= 1;
    }
  } else {
    __syncthreads();  // the offset tables
  }
  // epilogue: lane holds column lr of each 16 x 16 block, rows Acc<T>::row(lk, i)
  const int64_t cbase = lvl_off(a.Cb, b);
  if (a.nsplit > 1) {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int64_t n = n0 + nw + 16 * j + lr;
      if (n >= a.N) continue;
#pragma unroll
      for (int i = 0; i < MB; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int64_t m = m0 + mw + 16 * i + Acc<T>::row(lk, r);
          if (m < a.M) ((T*)a.partial)[(((int64_t)sp * a.batch + b) * a.M + m) * a.N + n] = acc[i][j][r];
        }
    }
    return;
  }
  // Each wave's 16-row slab of the tile goes through LDS so that lanes then
  // read and write C in runs of 4 consecutive columns (16 consecutive lanes
  // cover a slab row): coalesced, and 16-byte accesses where C allows.
  T(*sE)[16][EP] = (T(*)[16][EP])smem;
  constexpr int CH = 16 * WN / 4;  // 4-column chunks of a slab
#pragma unroll
  for (int i = 0; i < MB; i++) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) sE[w][Acc<T>::row(lk, r)][16 * j + lr] = acc[i][j][r];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < (CH + 63) / 64; q++) {
      const int c = lane + 64 * q;
      if (c >= CH) break;
      const int row = c / (WN / 4), col = (c % (WN / 4)) * 4;
      const int ml = mw + 16 * i + row;
      if (m0 + ml >= a.M) continue;
      const int nl = nw + col;
      const int64_t rowoff = cbase + cmoff[ml];
      double o[4];
#pragma unroll
      for (int e = 0; e < 4; e++) o[e] = a.alpha * (double)sE[w][row][col + e];
      if (a.c_vec && n0 + nl < a.N) {
        const int64_t off = rowoff + cnoff[nl];
        if (a.c64) {
          double2* p = (double2*)((double*)a.C + off);
          if (a.beta != 0.0) {
            const double2* q = (const double2*)((const double*)a.Csrc + off);
            const double2 x = q[0], y = q[1];
            o[0] += a.beta * x.x;
            o[1] += a.beta * x.y;
            o[2] += a.beta * y.x;
            o[3] += a.beta * y.y;
          }
          p[0] = make_double2(o[0], o[1]);
          p[1] = make_double2(o[2], o[3]);
        } else {
          float4* p = (float4*)((float*)a.C + off);
          if (a.beta != 0.0) {
            const float4 x = *(const float4*)((const float*)a.Csrc + off);
            o[0] += a.beta * x.x;
            o[1] += a.beta * x.y;
            o[2] += a.beta * x.z;
            o[3] += a.beta * x.w;
          }
          *p = make_float4((float)o[0], (float)o[1], (float)o[2], (float)o[3]);
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++) {
          if (n0 + nl + e >= a.N) break;
          const int64_t off = rowoff + cnoff[nl + e];
          if (a.beta != 0.0)
            o[e] += a.beta * (a.c64 ? ((const double*)a.Csrc)[off] : (double)((const float*)a.Csrc)[off]);
          if (a.c64)
            ((double*)a.C)[off] = o[e];
          else
            ((float*)a.C)[off] = (float)o[e];
        }
      }
    }
  }
}

// First level of a long split-K sum over small outputs: group g of PSG
// consecutive splits summed in split order, for every element (grid: element
// blocks x groups), into out[g][e].
constexpr int PSG = 32;
template <typename T>
__global__ void k_dgemm_reduce_groups(const T* __restrict__ partial, int nsplit, int64_t total,
                                      T* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int q0 = blockIdx.y * PSG, q1 = min(nsplit, q0 + PSG);
  T s = 0;
  for (int q = q0; q < q1; q++) s += partial[(int64_t)q * total + e];
  out[(int64_t)blockIdx.y * total + e] = s;
}

// C = alpha * sum over splits (in split order) + beta * C
template <typename T>
__global__ void k_dgemm_reduce(DArgs a) {
  const int64_t total = a.batch * a.M * a.N;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    T s = 0;
    for (int q = 0; q < a.nsplit; q++) s += ((const T*)a.partial)[(int64_t)q * total + e];
    const int64_t n = e % a.N, m = (e / a.N) % a.M, b = e / (a.N * a.M);
    const int64_t off = lvl_off(a.Cb, b) + lvl_off(a.Cm, m) + lvl_off(a.Cn, n);
    double o = a.alpha * (double)s;
    if (a.beta != 0.0) o += a.beta * (a.c64 ? ((const double*)a.Csrc)[off] : (double)((const float*)a.Csrc)[off]);
    if (a.c64)
      ((double*)a.C)[off] = o;
    else
      ((float*)a.C)[off] = (float)o;
  }
}

// ---------------------------------------------------------------------------
// Planner: einsum spec of two operands -> one strided batched GEMM.

struct Letter {
  char c;
  int64_t n;
  int64_t s[3];  // stride in A, B, C (0 where absent)
  bool in[3];
};

// Letters of a group, outer to inner, merged where strides compose in every
// operand that holds the group; at most three levels remain.  The order tried
// first is that of the larger operand (its innermost level then has the unit
// stride the tile loads walk), then the others.
static int group_levels(const std::vector<Letter>& g, const std::vector<int>& ops, std::vector<Letter>* out,
                        const char* what) {
  std::vector<Letter> v;
  for (auto& l : g)
    if (l.n > 1) v.push_back(l);
  std::vector<Letter> best;
  bool have = false;
  for (int op : ops) {
    std::vector<Letter> o = v;
    std::stable_sort(o.begin(), o.end(), [&](const Letter& x, const Letter& y) { return x.s[op] > y.s[op]; });
    std::vector<Letter> r;
    for (auto& l : o) {
      if (!r.empty()) {
        Letter& p = r.back();
        bool ok = true;
        for (int q : ops) ok = ok && p.s[q] == l.n * l.s[q];
        if (ok) {
          p.n *= l.n;
          for (int q = 0; q < 3; q++) p.s[q] = l.s[q];
          continue;
        }
      }
      r.push_back(l);
    }
    if (!have || r.size() < best.size()) {
      best = r;
      have = true;
    }
    if (best.size() <= 2) break;
  }
  FRICC_CHECK(best.size() <= 3, std::string("dense_contract: the ") + what +
                                    " letters do not reduce to three strided levels (copy an operand first)");
  *out = best;
  return OK;
}

// levels [outer .. inner] -> Lvl for operand op, and the group's extent
static Lvl make_lvl(const std::vector<Letter>& lv, int op, int64_t* extent) {
  Lvl d;
  d.nl = (int)lv.size();
  d.L0 = d.L1 = 1;
  d.s0 = d.s1 = d.s2 = 0;
  int64_t n = 1;
  const int k = (int)lv.size();
  if (k >= 1) {
    d.L0 = lv[k - 1].n;
    d.s0 = lv[k - 1].s[op];
  }
  if (k >= 2) {
    d.L1 = lv[k - 2].n;
    d.s1 = lv[k - 2].s[op];
  }
  if (k >= 3) d.s2 = lv[0].s[op];
  for (auto& l : lv) n *= l.n;
  d.r0 = 1.0 / (double)d.L0;
  d.r1 = 1.0 / (double)d.L1;
  *extent = n;
  return d;
}

template <typename T, int BM, int BN>
static void launch_tile(const DArgs& a, int va, int vb, hipStream_t s) {
  dim3 grid((unsigned)(a.ntm * a.ntn), (unsigned)(a.batch * a.nsplit));
  if (va == 4 && vb == 4)
    k_dgemm<T, BM, BN, 4, 4><<<grid, 256, 0, s>>>(a);
  else if (va == 4)
    k_dgemm<T, BM, BN, 4, 1><<<grid, 256, 0, s>>>(a);
  else if (vb == 4)
    k_dgemm<T, BM, BN, 1, 4><<<grid, 256, 0, s>>>(a);
  else
    k_dgemm<T, BM, BN, 1, 1><<<grid, 256, 0, s>>>(a);
}

template <typename T>
static int launch_dgemm(DArgs a, int bm, int bn, int va, int vb, hipStream_t s) {
#define FRICC_TILE(X, Y)                \
  if (bm == X && bn == Y) {             \
    launch_tile<T, X, Y>(a, va, vb, s); \
    return OK;                          \
  }
  FRICC_TILE(128, 128)
  FRICC_TILE(128, 64)
  FRICC_TILE(64, 128)
  FRICC_TILE(64, 64)
  FRICC_TILE(128, 32)
  FRICC_TILE(32, 128)
  FRICC_TILE(64, 32)
  FRICC_TILE(32, 64)
  FRICC_TILE(32, 32)
#undef FRICC_TILE
  set_error("dense_contract: no kernel for this tile shape");
  return EINVAL_;
}

static int pick_tile(int64_t n) { return n <= 32 ? 32 : n <= 64 ? 64 : 128; }

int dense_contract(const char* spec, double alpha, const void* A, int a64, int an, const int64_t* ash,
                   const int64_t* ast, const void* B, int b64, int bn, const int64_t* bsh, const int64_t* bst,
                   double beta, void* C, int c64, int cn, const int64_t* csh, const int64_t* cst, int compute64,
                   hipStream_t s, const void* Csrc) {
  if (!Csrc) Csrc = C;
  FRICC_CHECK(spec && A && B && C, "dense_contract: null argument");
  FRICC_CHECK(an >= 0 && an <= 8 && bn >= 0 && bn <= 8 && cn >= 0 && cn <= 8, "dense_contract: at most 8 axes");
  const char* comma = strchr(spec, ',');
  const char* arrow = strstr(spec, "->");
  FRICC_CHECK(comma && arrow && comma < arrow, "dense_contract: spec must look like 'ab,bc->ac'");
  const std::string sa(spec, comma), sb(comma + 1, arrow), sc(arrow + 2);
  FRICC_CHECK((int)sa.size() == an && (int)sb.size() == bn && (int)sc.size() == cn,
              "dense_contract: spec and tensor ranks differ");
  std::vector<Letter> lets;
  auto add = [&](const std::string& str, const int64_t* sh, const int64_t* st, int op) -> int {
    for (size_t i = 0; i < str.size(); i++) {
      const char c = str[i];
      FRICC_CHECK(c >= 'a' && c <= 'z', "dense_contract: letters must be a-z");
      for (size_t j = 0; j < i; j++) FRICC_CHECK(str[j] != c, "dense_contract: repeated letter in one operand");
      Letter* l = nullptr;
      for (auto& x : lets)
        if (x.c == c) l = &x;
      if (!l) {
        Letter nl;
        nl.c = c;
        nl.n = sh[i];
        for (int q = 0; q < 3; q++) {
          nl.s[q] = 0;
          nl.in[q] = false;
        }
        lets.push_back(nl);
        l = &lets.back();
      }
      FRICC_CHECK(sh[i] >= 1, "dense_contract: empty axis");
      FRICC_CHECK(l->n == sh[i], std::string("dense_contract: extents of letter ") + c + " differ");
      l->s[op] = st[i];
      l->in[op] = true;
    }
    return OK;
  };
  FRICC_TRY(add(sa, ash, ast, 0));
  FRICC_TRY(add(sb, bsh, bst, 1));
  FRICC_TRY(add(sc, csh, cst, 2));
  std::vector<Letter> gb, gm, gn, gk;
  for (auto& l : lets) {
    const bool ia = l.in[0], ib = l.in[1], ic = l.in[2];
    if (ia && ib && ic)
      gb.push_back(l);
    else if (ia && ic)
      gm.push_back(l);
    else if (ib && ic)
      gn.push_back(l);
    else if (ia && ib)
      gk.push_back(l);
    else
      FRICC_CHECK(l.n == 1, std::string("dense_contract: letter ") + l.c + " appears in only one operand");
  }
  std::vector<Letter> lb, lm, ln, lk;
  int64_t sizes[3] = {1, 1, 1};
  for (int i = 0; i < an; i++) sizes[0] *= ash[i];
  for (int i = 0; i < bn; i++) sizes[1] *= bsh[i];
  for (int i = 0; i < cn; i++) sizes[2] *= csh[i];
  auto by_size = [&](std::vector<int> ops) {
    std::stable_sort(ops.begin(), ops.end(), [&](int x, int y) { return sizes[x] > sizes[y]; });
    return ops;
  };
  FRICC_TRY(group_levels(gb, by_size({0, 1, 2}), &lb, "batch"));
  FRICC_TRY(group_levels(gm, by_size({0, 2}), &lm, "m"));
  FRICC_TRY(group_levels(gn, by_size({1, 2}), &ln, "n"));
  FRICC_TRY(group_levels(gk, by_size({0, 1}), &lk, "k"));
  DArgs a;
  int64_t nb, M, N, K, t;
  a.Ab = make_lvl(lb, 0, &nb);
  a.Bb = make_lvl(lb, 1, &t);
  a.Cb = make_lvl(lb, 2, &t);
  a.Am = make_lvl(lm, 0, &M);
  a.Cm = make_lvl(lm, 2, &t);
  a.Bn = make_lvl(ln, 1, &N);
  a.Cn = make_lvl(ln, 2, &t);
  a.Ak = make_lvl(lk, 0, &K);
  a.Bk = make_lvl(lk, 1, &t);
  FRICC_CHECK(nb < ((int64_t)1 << 31) && M < ((int64_t)1 << 31) && N < ((int64_t)1 << 31) && K < ((int64_t)1 << 31),
              "dense_contract: an index group exceeds 2^31");
  a.A = A;
  a.B = B;
  a.C = C;
  a.Csrc = Csrc;
  a.a64 = a64;
  a.b64 = b64;
  a.c64 = c64;
  a.M = M;
  a.N = N;
  a.K = K;
  a.batch = nb;
  a.alpha = alpha;
  a.beta = beta;
  // coalescing: walk the unit-stride index of each tile
  a.a_kfast = (a.Ak.s0 == 1 && a.Am.s0 != 1) ? 1 : 0;
  a.b_kfast = (a.Bk.s0 == 1 && a.Bn.s0 != 1) ? 1 : 0;
  int bm = pick_tile(M);
  const int bnt = pick_tile(N);
  // FP64 128 x 128 tiles hold 128 accumulator registers per lane and run one
  // wave per SIMD; 64 x 128 tiles run two
  if (compute64 && bm == 128 && bnt == 128) bm = 64;
  a.ntm = ceil_div(M, bm);
  a.ntn = ceil_div(N, bnt);
  // 16-byte loads along the unit-stride index of an operand when every group
  // of 4 elements there stays in one level and starts 16-byte aligned
  auto mult4 = [](const Lvl& d) { return d.s1 % 4 == 0 && d.s2 % 4 == 0; };
  auto vec_ok = [&](const void* p, const Lvl& fast, const Lvl& other, const Lvl& bat) {
    return fast.s0 == 1 && fast.L0 % 4 == 0 && mult4(fast) && other.s0 % 4 == 0 && mult4(other) &&
           bat.s0 % 4 == 0 && mult4(bat) && ((uintptr_t)p % 16) == 0;
  };
  const int va = a.a_kfast ? (vec_ok(A, a.Ak, a.Am, a.Ab) ? 4 : 1) : (vec_ok(A, a.Am, a.Ak, a.Ab) ? 4 : 1);
  const int vb = a.b_kfast ? (vec_ok(B, a.Bk, a.Bn, a.Bb) ? 4 : 1) : (vec_ok(B, a.Bn, a.Bk, a.Bb) ? 4 : 1);
  a.c_vec = (a.Cn.s0 == 1 && a.Cn.L0 % 4 == 0 && mult4(a.Cn) && a.Cm.s0 % 4 == 0 && mult4(a.Cm) &&
             a.Cb.s0 % 4 == 0 && mult4(a.Cb) && ((uintptr_t)C % 16) == 0 && ((uintptr_t)Csrc % 16) == 0)
                ? 1
                : 0;
  // split K when the output grid cannot fill the device (memory-bound
  // products need several blocks per CU in flight)
  const int64_t tiles = a.ntm * a.ntn * nb;
  int nsplit = 1;
  if (tiles < 1024 && K >= 8 * DBK) {
    // about 2048 blocks; tiny outputs (a few tiles, a very long sum) may take
    // more splits, their partial sums being small
    const int64_t cap = M * N <= 4096 ? 1024 : 256;
    nsplit = (int)std::min<int64_t>(std::min<int64_t>(ceil_div(2048, tiles), K / (4 * DBK)), cap);
    nsplit = std::max(nsplit, 1);
  }
  a.kchunk = ceil_div(ceil_div(K, nsplit), DBK) * DBK;
  nsplit = (int)ceil_div(K, a.kchunk);
  a.nsplit = std::max(nsplit, 1);
  FRICC_CHECK(nb * a.nsplit < 65536 && a.ntm * a.ntn < ((int64_t)1 << 31), "dense_contract: grid too large");
  a.partial = nullptr;
  if (a.nsplit > 1) {
    void* p = nullptr;
    FRICC_TRY(ws_get(50, (size_t)a.nsplit * nb * M * N * (compute64 ? 8 : 4), &p));
    a.partial = p;
  }
  if (M == 0 || N == 0 || nb == 0) return OK;
  if (compute64)
    FRICC_TRY(launch_dgemm<double>(a, bm, bnt, va, vb, s));
  else
    FRICC_TRY(launch_dgemm<float>(a, bm, bnt, va, vb, s));
  if (a.nsplit > 1) {
    const int64_t total = nb * M * N;
    if (a.nsplit > 2 * PSG && ceil_div(total, 256) < 512) {
      // few outputs and many splits: one thread per element would sum
      // hundreds of partials on a handful of CUs; sum groups of PSG splits
      // in parallel first (a fixed order: deterministic)
      const int ng = (int)ceil_div(a.nsplit, PSG);
      void* p2 = nullptr;
      FRICC_TRY(ws_get(51, (size_t)ng * total * (compute64 ? 8 : 4), &p2));
      const dim3 gg((unsigned)ceil_div(total, 256), (unsigned)ng);
      if (compute64)
        k_dgemm_reduce_groups<double><<<gg, 256, 0, s>>>((const double*)a.partial, a.nsplit, total, (double*)p2);
      else
        k_dgemm_reduce_groups<float><<<gg, 256, 0, s>>>((const float*)a.partial, a.nsplit, total, (float*)p2);
      a.partial = p2;
      a.nsplit = ng;
    }
    const unsigned g = (unsigned)std::min<int64_t>(ceil_div(total, 256), 4096);
    if (compute64)
      k_dgemm_reduce<double><<<g, 256, 0, s>>>(a);
    else
      k_dgemm_reduce<float><<<g, 256, 0, s>>>(a);
  }
  FRICC_HIP(hipGetLastError());
  return OK;
}

// ---------------------------------------------------------------------------
// Two products that read the same o*v^3 operand Z[k,c,d,a], fused so that Z
// is streamed once instead of twice (T = float: FP32 operands and MFMA;
// T = double: FP64):
//   out_ia[i,a] += sum_{k,c,d} t2[i,k,c,d] Z[k,c,d,a]   (MFMA 16x16x4)
//   out_ac[a,c]  = sum_{k,d}   t1[k,d]     Z[k,c,d,a]   (FMA)
// A workgroup owns one c and 128 a's and walks K = (k, d) in steps of KD d
// (KD * sizeof(T) = 128 bytes: 32 floats or 16 doubles): per step it stages
// t2[0..31, k, c, d0..d0+KD), Z[k, c, d0..d0+KD, a0..+127] and
// t1[k, d0..d0+KD) through three LDS stages (global->LDS DMA, two steps in
// flight while one is computed).  The first product's i x a block goes to a
// per-c partial (fixed-order FP64 sum over c in k_zpair_reduce); the second
// is complete in the workgroup (its sum over k runs inside).  Needs o <= 32
// and v % 128 == 0; rows i >= o read row o-1 and are not written.
constexpr int ZP_BA = 128;  // a per workgroup
constexpr int ZP_S = 3;     // LDS stages
constexpr int ZP_LW = 6;    // DMA instructions per wave per step (A 1, B 4, t1 1)

template <typename T>
struct ZP {
  static constexpr int KD = 128 / (int)sizeof(T);          // d per step
  static constexpr int RPD = 1024 / (ZP_BA * (int)sizeof(T));  // Z rows per 1 KB DMA
  static constexpr int LPR = 64 / RPD;                     // lanes per Z row
};

template <typename T>
struct ZPStage {
  T A[32][ZP<T>::KD];     // t2 rows i
  T B[ZP<T>::KD][ZP_BA];  // Z rows d
  T t1[ZP<T>::KD];
};

template <typename T>
struct ZPArgs {
  const T* t2;
  const T* t1;
  const T* Z;
  T* part;  // [v][32][v]
  double* out_ac;
  int o, v;
};

// DMAs of step q (k = q / (v/KD), d0 = KD (q % (v/KD))) into one stage
template <typename T>
__device__ inline void zp_issue(const ZPArgs<T>& a, int c, int a0, int q, ZPStage<T>& st, int w, int lane) {
  constexpr int KD = ZP<T>::KD, RPD = ZP<T>::RPD, LPR = ZP<T>::LPR, E16 = 16 / (int)sizeof(T);
  const int nd = a.v / KD;
  const int k = q / nd, d0 = (q % nd) * KD;
  const int64_t v = a.v;
  // A: rows 8w..8w+7, 8 lanes of 16 bytes per 128-byte row
  {
    const int i = min(8 * w + (lane >> 3), a.o - 1);
    const T* g = a.t2 + (((int64_t)i * a.o + k) * v + c) * v + d0 + E16 * (lane & 7);
    dma16(g, lds_addr(&st.A[8 * w][0]));
  }
  // B: rows (KD/4)w .. +KD/4, RPD rows (LPR lanes x 16 bytes each) per DMA
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int r0 = (KD / 4) * w + RPD * j;
    const T* g = a.Z + (((int64_t)k * v + c) * v + d0 + r0 + lane / LPR) * v + a0 + E16 * (lane % LPR);
    dma16(g, lds_addr(&st.B[r0][0]));
  }
  // t1[k, d0..d0+KD): lanes 0-7 of every wave write the same 128 bytes
  if (lane < 8) dma16(a.t1 + (int64_t)k * v + d0 + E16 * lane, lds_addr(&st.t1[0]));
}

template <typename T>
__global__ __launch_bounds__(256, 2) void k_zpair(ZPArgs<T> a) {
  constexpr int KD = ZP<T>::KD;
  typedef typename Acc<T>::type V4;
  __shared__ ZPStage<T> st[ZP_S];
  __shared__ T red[ZP_BA];
  const int c = blockIdx.x / (a.v / ZP_BA);
  const int a0 = (blockIdx.x % (a.v / ZP_BA)) * ZP_BA;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int nq = a.o * (a.v / KD);
  V4 acc[2][2];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) acc[rb][cb] = V4{0, 0, 0, 0};
  T acc2 = 0;
  const int ac = tid & (ZP_BA - 1), h = tid >> 7;  // the second product: column ac, rows (KD/2)h ..
#pragma unroll
  for (int s = 0; s < ZP_S - 1; s++)
    if (s < nq) zp_issue(a, c, a0, s, st[s], w, lane);
  int cur = 0;
  for (int q = 0; q < nq; q++) {
    if (q + 1 < nq)
      wait_vm<ZP_LW>();
    else
      wait_vm<0>();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    const int nxt = cur == 0 ? ZP_S - 1 : cur - 1;  // (cur + S - 1) % S
    if (q + ZP_S - 1 < nq) zp_issue(a, c, a0, q + ZP_S - 1, st[nxt], w, lane);
    const ZPStage<T>& sm = st[cur];
#pragma unroll
    for (int s = 0; s < KD / 4; s++) {
      const T x0 = sm.A[lr][4 * s + lk], x1 = sm.A[16 + lr][4 * s + lk];
#pragma unroll
      for (int cb = 0; cb < 2; cb++) {
        const T y = sm.B[4 * s + lk][w * 32 + cb * 16 + lr];
        acc[0][cb] = Acc<T>::mfma(x0, y, acc[0][cb]);
        acc[1][cb] = Acc<T>::mfma(x1, y, acc[1][cb]);
      }
    }
#pragma unroll
    for (int j = 0; j < KD / 2; j++) acc2 = fma(sm.t1[(KD / 2) * h + j], sm.B[(KD / 2) * h + j][ac], acc2);
    cur = cur == ZP_S - 1 ? 0 : cur + 1;
  }
  // first product: this c's partial of rows i < o
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
      for (int ii = 0; ii < 4; ii++) {
        const int i = rb * 16 + Acc<T>::row(lk, ii);
        if (i < a.o) a.part[((int64_t)c * 32 + i) * a.v + a0 + w * 32 + cb * 16 + lr] = acc[rb][cb][ii];
      }
  // second product: add the two row halves, write out_ac[a][c]
  if (h == 1) red[ac] = acc2;
  __syncthreads();
  if (h == 0) a.out_ac[(int64_t)(a0 + ac) * a.v + c] = (double)acc2 + (double)red[ac];
}

// out_ia[i,a] += sum over c of the partials, in c order, in FP64
template <typename T>
__global__ void k_zpair_reduce(const T* __restrict__ part, int o, int v, double* __restrict__ out_ia) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)o * v) return;
  const int i = (int)(e / v), aa = (int)(e % v);
  double s = 0.0;
  for (int c = 0; c < v; c++) s += (double)part[((int64_t)c * 32 + i) * v + aa];
  out_ia[e] += s;
}

template <typename T>
static int zpair_launch(int o, int v, const T* t2, const T* t1, const T* Z, double* out_ia, double* out_ac,
                        hipStream_t s) {
  void* p;
  FRICC_TRY(ws_get(52, (size_t)v * 32 * v * sizeof(T), &p));
  ZPArgs<T> a;
  a.t2 = t2;
  a.t1 = t1;
  a.Z = Z;
  a.part = (T*)p;
  a.out_ac = out_ac;
  a.o = o;
  a.v = v;
  k_zpair<T><<<v * (v / ZP_BA), 256, 0, s>>>(a);
  FRICC_HIP(hipGetLastError());
  k_zpair_reduce<T><<<(unsigned)ceil_div((int64_t)o * v, 256), 256, 0, s>>>((const T*)p, o, v, out_ia);
  FRICC_HIP(hipGetLastError());
  return OK;
}

int dense_zpair(int o, int v, int f64, const void* t2, const void* t1, const void* Z, double* out_ia,
                double* out_ac, hipStream_t s) {
  FRICC_CHECK(o > 0 && o <= 32 && v > 0 && v % ZP_BA == 0, "dense_zpair: needs 0 < o <= 32 and v a multiple of 128");
  FRICC_CHECK(t2 && t1 && Z && out_ia && out_ac, "dense_zpair: null pointer");
  FRICC_CHECK(((uintptr_t)t2 | (uintptr_t)t1 | (uintptr_t)Z) % 16 == 0, "dense_zpair: operands must be 16-byte aligned");
  if (f64)
    return zpair_launch<double>(o, v, (const double*)t2, (const double*)t1, (const double*)Z, out_ia, out_ac, s);
  return zpair_launch<float>(o, v, (const float*)t2, (const float*)t1, (const float*)Z, out_ia, out_ac, s);
}

}  // namespace fricc

extern "C" {
// C = alpha * einsum(spec, A, B) + beta * C for two strided device tensors
// (see the top of this file).  *_precision: 0 FP64, 1 FP32 element types;
// compute_precision: 0 FP64 MFMA, 1 FP32 MFMA.  No counterpart in the
// reference (its dense products are PySCF/NumPy einsums,
// fricc/friccsd.py:144-207, 309-376).
int fricc_dense_contract(const char* spec, double alpha, const void* A, int a_precision, int a_ndim,
                         const int64_t* a_shape, const int64_t* a_strides, const void* B, int b_precision, int b_ndim,
                         const int64_t* b_shape, const int64_t* b_strides, double beta, void* C, int c_precision,
                         int c_ndim, const int64_t* c_shape, const int64_t* c_strides, int compute_precision,
                         void* stream) {
  FRICC_CHECK((a_precision == 0 || a_precision == 1) && (b_precision == 0 || b_precision == 1) &&
                  (c_precision == 0 || c_precision == 1) && (compute_precision == 0 || compute_precision == 1),
              "dense_contract: precisions must be 0 (FP64) or 1 (FP32)");
  return fricc::dense_contract(spec, alpha, A, a_precision == 0, a_ndim, a_shape, a_strides, B, b_precision == 0,
                               b_ndim, b_shape, b_strides, beta, C, c_precision == 0, c_ndim, c_shape, c_strides,
                               compute_precision == 0, (hipStream_t)stream, nullptr);
}

// C = alpha * einsum(spec, A, B) + beta * Csrc, Csrc of C's shape, strides and
// element type (C itself, or a tensor that does not overlap it): an
// accumulation onto a constant operand without copying it into C first.
// No counterpart in the reference.
int fricc_dense_contract_from(const char* spec, double alpha, const void* A, int a_precision, int a_ndim,
                              const int64_t* a_shape, const int64_t* a_strides, const void* B, int b_precision,
                              int b_ndim, const int64_t* b_shape, const int64_t* b_strides, double beta,
                              const void* Csrc, void* C, int c_precision, int c_ndim, const int64_t* c_shape,
                              const int64_t* c_strides, int compute_precision, void* stream) {
  FRICC_CHECK((a_precision == 0 || a_precision == 1) && (b_precision == 0 || b_precision == 1) &&
                  (c_precision == 0 || c_precision == 1) && (compute_precision == 0 || compute_precision == 1),
              "dense_contract: precisions must be 0 (FP64) or 1 (FP32)");
  FRICC_CHECK(Csrc, "dense_contract_from: null source");
  return fricc::dense_contract(spec, alpha, A, a_precision == 0, a_ndim, a_shape, a_strides, B, b_precision == 0,
                               b_ndim, b_shape, b_strides, beta, C, c_precision == 0, c_ndim, c_shape, c_strides,
                               compute_precision == 0, (hipStream_t)stream, Csrc);
}

// out_ia[i,a] += sum_{k,c,d} t2[i,k,c,d] Z[k,c,d,a] and out_ac[a,c] =
// sum_{k,d} t1[k,d] Z[k,c,d,a] with Z read once (operands FP64 with precision
// 0 or FP32 with precision 1, on the MFMA of that type; FP64 outputs).
// With Z[k,c,d,a] = 2 (kd|ac) - (kc|ad) these are the t1 equation's ovvv*t2
// terms (fricc/friccsd.py:165-166) and the ovvv*t1 part of Lvv
// (imd.Lvv, fricc/friccsd.py:190; also the t1*t1 terms of :167-168), which
// the reference evaluates as separate einsums; the fused entry point itself
// has no counterpart in the reference.
int fricc_dense_zpair(int o, int v, int precision, const void* t2, const void* t1, const void* Z, double* out_ia,
                      double* out_ac, void* stream) {
  FRICC_CHECK(precision == 0 || precision == 1, "dense_zpair: precision must be 0 (FP64) or 1 (FP32)");
  return fricc::dense_zpair(o, v, precision == 0, t2, t1, Z, out_ia, out_ac, (hipStream_t)stream);
}
}
