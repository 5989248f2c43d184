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
#include "mfma.h"
#include "dma.h"

#include <utility>

namespace fricc {

struct OutDev {
  void* out;
  int of32;
  uint32_t Xl, Yl;
  int64_t oxh, oxl, oyh, oyl;
  int vec;  // runs of 4 consecutive elements along y (1) or x (2) are contiguous and 16-byte aligned; 0: neither
};

// out[off] += alpha * v in the output's element type
__device__ inline void out_add(const OutDev& om, int64_t off, double alpha, double v) {
  if (om.of32)
    ((float*)om.out)[off] += (float)(alpha * v);
  else
    ((double*)om.out)[off] += alpha * v;
}

__device__ inline int64_t out_xoff(const OutDev& om, int64_t x) {
  const uint32_t xx = (uint32_t)x;
  return (int64_t)(xx / om.Xl) * om.oxh + (int64_t)(xx % om.Xl) * om.oxl;
}

// out[off[n]] += alpha * v[n] for the N outputs of one lane and column block
// (ok[n] false: skipped).  All N values are read before any is written, so the
// round trips overlap; one read-modify-write at a time would serialise them,
// since the compiler cannot tell that the offsets differ.
template <typename TO, int N>
__device__ inline void out_add_batch(TO* __restrict__ out, const int64_t (&off)[N], const bool (&ok)[N], double alpha,
                                     const double (&v)[N]) {
  TO cur[N];
#pragma unroll
  for (int n = 0; n < N; n++) cur[n] = ok[n] ? out[off[n]] : (TO)0;
#pragma unroll
  for (int n = 0; n < N; n++)
    if (ok[n]) out[off[n]] = cur[n] + (TO)(alpha * v[n]);
}

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
    int64_t off[16];
    bool ok[16];
    double vv[16];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int64_t x = xw + 16 * s + Acc<T>::row(lk, i);
        ok[4 * s + i] = x < X;
        off[4 * s + i] = out_xoff(om, x < X ? x : 0) + oy;
        vv[4 * s + i] = (double)acc[s][q][i];
      }
    if (om.of32)
      out_add_batch<float, 16>((float*)om.out, off, ok, alpha, vv);
    else
      out_add_batch<double, 16>((double*)om.out, off, ok, alpha, vv);
  }
}

int g_order = 1;
int g_split = 0;   // forced K split of the BF16 gathered GEMM (0: automatic)
int g_fp_lds = 1;  // FP64/FP32 through the LDS pipeline where the geometry allows


// ---------------------------------------------------------------------------
// BF16 gathered contraction on v_mfma_f32_16x16x32_bf16 (FP32 accumulate).
//
// A comes as 1 KB records [X/256][R][hi 256 | lo 256] bf16 (X rounded up to
// a multiple of 256, pad columns zero): the block of columns one workgroup
// reads of row r is one contiguous record, and the rows of one column block
// lie together, so the rows a K step gathers stay close in memory.  B (the
// view) comes as hi/lo bf16 planes [k/8][TY][8] over the tile's gathered
// rows k.  npass 3: hi*hi + hi*lo + lo*hi (about 16
// significant bits per product), npass 1: hi*hi.
//
// Rows are gathered one by one (not in fixed groups of consecutive rows), so
// a tile's K range holds only rows that carry entries of that tile.  Per K
// step the block stages the 32 gathered rows of A, columns [x0, +BX), as
// row-major LDS rows (one 1 KB DMA per row: hi then lo), and B's four
// 8-row groups.  The MFMA A fragment wants 8 consecutive k per lane for one
// x; ds_read_b64_tr_b16 delivers exactly that from k-major LDS rows: within
// each 16-lane group, lane t receives element t%4 of the 8-byte chunks that
// lanes t/4, t/4+4, t/4+8, t/4+12 address, so lane s addressing row
// k0 + s/4, columns 4*(s%4).. hands lane t the column t of rows k0..k0+3.
// Two such reads give rows k0..k0+7.
//
// A workgroup (4 waves) owns a BX x BY output block: x in [bx*BX, +BX), y a
// BY-wide slice of one view tile; waves are 2 (x) x 2 (y), each owning
// 128 x 64 = 8 x 4 MFMA blocks.  Consecutive blocks on one XCD are the BY
// slices of one tile (same gathered A rows, shared through that XCD's L2).
typedef __attribute__((ext_vector_type(8))) __bf16 bf8;
typedef short v4s __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v4s lds_v4s;
typedef __attribute__((address_space(3))) uint8_t lds_u8;

__device__ inline bf8 as_bf8(uint4 u) {
  bf8 r;
  __builtin_memcpy(&r, &u, 16);
  return r;
}

typedef short v8s __attribute__((ext_vector_type(8)));

__device__ inline bf8 cat_bf8(v4s a, v4s b) {
  const v8s c = __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf8, c);
}

constexpr int BX = 256;
constexpr int KG = 4;                  // 8-row groups of B per K step
constexpr int KR = 32;                 // gathered rows per K step (one MFMA K)

// Geometry of the staged A rows.  HILO: the operand's records carry the lo
// plane (1 KB: hi 256 | lo 256); else only the hi plane (512 B), the operand
// of single-pass products.
template <bool HILO>
struct AGeom {
  static constexpr int REC = HILO ? 2 * BX : BX;  // bf16 per record
  static constexpr int ROW = REC * 2;             // staged bytes of one A row
  static constexpr int PITCH = ROW + 32;          // LDS pitch of a staged A row
  // byte offset of staged row k: rows k and k+8 would meet the same banks at
  // this pitch, so each group of 8 rows starts 128 bytes (half the banks)
  // further on than the pitch alone would place it
  static constexpr int GROUP = 8 * PITCH + 128;
  static constexpr int STAGE = (KR / 8) * GROUP;
  static __device__ constexpr uint32_t off(int k) { return (uint32_t)((k >> 3) * GROUP + (k & 7) * PITCH); }
  static_assert(7 * PITCH + ROW <= GROUP, "staged rows must not overlap");
};

struct B16Args {
  const uint16_t* A;  // [X/256][R][hi 256 | lo 256], or [X/256][R][hi 256] for hi-only operands
  int64_t R, X, Y;
  const uint32_t* tile_off;
  const uint32_t* tile_order;  // tiles by decreasing row count
  const uint32_t* glist;
  const uint4* Bhi;  // view planes [k/8][TY][8] bf16
  const uint4* Blo;
  int64_t TY;
  int64_t nbx, nyb;
  int nsplit, order;
  OutDev om;
  double alpha;
  float* partial;  // [nsplit][X][Y] when nsplit > 1
};

template <int BYB, bool HILO>
struct StageMem {
  uint8_t A[AGeom<HILO>::STAGE];  // staged rows [k]: hi BX (| lo BX) bf16, at AGeom::off(k)
  uint4 B[2][KG][BYB];            // [hi|lo][group][y]
};

// Operands of the pipeline, copied out of the kernel arguments once.  The
// row list is read through the constant address space: it is never written
// while the kernel runs, so its loads stay scalar (no vector-memory wait).
typedef const __attribute__((address_space(4))) uint32_t cu32;
struct StageSrc {
  const uint16_t* A;  // this lane's 16 bytes of record (column block, row 0)
  const uint4* Bhi;
  const uint4* Blo;
  cu32* glist;  // row list (constant address space)
  int64_t TY;
};

// One pipeline stage: gathered rows [k, k+32) of the tile, the A record of
// the workgroup's column block and B columns [bcol, +BYB), hi and lo planes.
// Wave w stages A rows 8w..8w+7 (one 1 KB record per DMA: lanes 0-31 the hi
// half, 32-63 the lo half) and a quarter of the B chunks (64 columns of one
// 8-row group and plane per DMA).
// The 8 gathered row indices wave w stages at K step k.  Inside the
// pipeline they come from LDS: one iteration before a stage's DMAs are
// issued, lanes 0-1 of each wave DMA that stage's 8 indices (32 bytes) from
// the row list into the wave's slot of a two-slot LDS table (row_dma), in
// issue order with the A/B DMAs, so the wait at the top of the next
// iteration also covers them; the wave then reads them with one LDS read
// before the barrier.  A scalar load here (row_scalar, used by the prologue
// only) would put its round trip between the barrier and the MFMAs, and a
// vector load into registers would make the compiler wait for the stage's
// DMAs before the MFMAs (it cannot see that the DMAs were issued later).
struct RowIdx {
  uint4 lo, hi;
};

__device__ inline RowIdx row_scalar(const StageSrc& p, uint32_t k, int w) {
  cu32* rl = p.glist + k + 8 * w;
  RowIdx r;
  r.lo = uint4{rl[0], rl[1], rl[2], rl[3]};
  r.hi = uint4{rl[4], rl[5], rl[6], rl[7]};
  return r;
}

// LDS table of prefetched row indices: [slot][wave] 8 indices
struct RowTab {
  uint4 r[2][4][2];
};

__device__ inline void row_dma(const StageSrc& p, uint32_t k, int w, int lane, RowTab& tab, int slot) {
  // k + 8w is a multiple of 8 and the row list is 16-byte aligned: lane l
  // (0 or 1) moves indices 4l..4l+3, 16 bytes, to tab.r[slot][w][l]
  if (lane < 2) dma16((const uint32_t*)p.glist + k + 8 * w + 4 * lane, lds_addr(&tab.r[slot][w][0]));
}

__device__ inline RowIdx row_lds(const RowTab& tab, int slot, int w) {
  RowIdx r;
  r.lo = tab.r[slot][w][0];
  r.hi = tab.r[slot][w][1];
  return r;
}

template <int NPASS, int BYB, bool HILO>
__device__ inline void stage_issue(const StageSrc& p, uint32_t k, int64_t bcol, StageMem<BYB, HILO>& sm, int w,
                                   int lane, const RowIdx& ri) {
  static_assert(BX == 256 && (BYB == 128 || BYB == 256) && KG == 4 && KR == 32,
                "wave assignment assumes 4 waves, BX 256, BYB 128 or 256");
  typedef AGeom<HILO> G;
  const uint32_t la = lds_addr(&sm.A[0]), lb = lds_addr(&sm.B[0][0][0]);
  const uint32_t rr[8] = {ri.lo.x, ri.lo.y, ri.lo.z, ri.lo.w, ri.hi.x, ri.hi.y, ri.hi.z, ri.hi.w};
  // one record per DMA: 64 lanes for a hi | lo record, lanes 0-31 for a hi one
#pragma unroll
  for (int i = 0; i < 8; i++)
    if (HILO || lane < 32) dma16(p.A + (int64_t)rr[i] * G::REC, la + G::off(8 * w + i));
  constexpr int NCH = BYB / 64;          // 64-column chunks per group
  constexpr int PER = KG * NCH / 4;      // chunks per wave
  const uint32_t kg = k >> 3;
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int c = w * PER + j;
    const int gl = c / NCH, yc = c % NCH;
    const int64_t off = (int64_t)(kg + gl) * p.TY + bcol + yc * 64 + lane;
    const uint32_t l = lb + (uint32_t)((gl * BYB + yc * 64) * 16);
    dma16(p.Bhi + off, l);
    if (NPASS > 1) dma16(p.Blo + off, l + (uint32_t)(KG * BYB * 16));
  }
}

template <int BYB, bool HILO>
__device__ inline v4s tr_read(const StageMem<BYB, HILO>& sm, uint32_t off) {
  return __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s*)((lds_u8*)(lds_void*)&sm.A[0] + off));
}

// MFMAs of one stage: wave (wx, wy) owns x [wx*128, +128) and y [wy*BYB/2, +BYB/2)
template <int NPASS, int BYB, bool HILO>
__device__ inline void stage_compute(const StageMem<BYB, HILO>& sm, f4 (&acc)[8][BYB / 32], int wx, int wy, int lr,
                                     int lk) {
  static_assert(HILO || NPASS == 1, "the lo plane of A is needed for three passes");
  typedef AGeom<HILO> G;
  constexpr int NQ = BYB / 32;
  uint4 bh[NQ], bl[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    bh[q] = sm.B[0][lk][wy * (BYB / 2) + 16 * q + lr];
    if (NPASS > 1) bl[q] = sm.B[1][lk][wy * (BYB / 2) + 16 * q + lr];
  }
  // this lane's transposed-read address: row 8*lk + lr/4 (+4 for the second
  // read), columns 4*(lr%4).. of each 16-column block
  const uint32_t r0 = G::off(8 * lk + (lr >> 2)) + (uint32_t)((lr & 3) * 8);
  const uint32_t r1 = G::off(8 * lk + 4 + (lr >> 2)) + (uint32_t)((lr & 3) * 8);
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const uint32_t c = (uint32_t)((wx * 128 + 16 * s) * 2);
    const bf8 ah = cat_bf8(tr_read(sm, r0 + c), tr_read(sm, r1 + c));
    bf8 al;
    if (NPASS > 1) al = cat_bf8(tr_read(sm, r0 + BX * 2 + c), tr_read(sm, r1 + BX * 2 + c));
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      acc[s][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, as_bf8(bh[q]), acc[s][q], 0, 0, 0);
      if (NPASS > 1) {
        acc[s][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, as_bf8(bl[q]), acc[s][q], 0, 0, 0);
        acc[s][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, as_bf8(bh[q]), acc[s][q], 0, 0, 0);
      }
    }
  }
}

// A workgroup of 4 waves owns a BX x BYB output block: x in [bx*BX, +BX), y a
// BYB-wide slice of one view tile; each wave owns 128 x BYB/2.  K advances 32
// gathered rows (one MFMA K) per stage; stages are filled by global->LDS DMA
// S-1 stages ahead of the MFMAs that read them, so no staging registers are
// needed and every element is fetched once per workgroup.  Each stage buffer
// is a separate LDS variable so in-flight DMA provably does not alias the
// stage being read.  Consecutive workgroups on one XCD are the BYB slices of
// one tile (same gathered A rows, shared through that XCD's L2).
// Epilogue of one 16-column block q of a wave's accumulators (compile-time q,
// so the accumulators stay in registers): split-K partials, or
// out += alpha * acc in two batches of 16 read-modify-writes.
template <int Q, int NQ>
__device__ inline void bf16_epilogue_q(const B16Args& a, const f4 (&acc)[8][NQ], int64_t x0, int64_t y0, int wx, int wy,
                                       int lr, int lk, int sp, int byb) {
  const int64_t y = y0 + wy * (byb / 2) + 16 * Q + lr;
  if (y >= a.Y) return;
  if (a.nsplit > 1) {
#pragma unroll
    for (int s = 0; s < 8; s++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int64_t x = x0 + wx * 128 + 16 * s + 4 * lk + i;
        if (x < a.X) a.partial[((int64_t)sp * a.X + x) * a.Y + y] = acc[s][Q][i];
      }
    return;
  }
  const uint32_t yy = (uint32_t)y;
  const int64_t oy = (int64_t)(yy / a.om.Yl) * a.om.oyh + (int64_t)(yy % a.om.Yl) * a.om.oyl;
#pragma unroll
  for (int hh = 0; hh < 2; hh++) {
    int64_t off[16];
    bool ok[16];
    double vv[16];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int64_t x = x0 + wx * 128 + 16 * (4 * hh + s) + 4 * lk + i;
        ok[4 * s + i] = x < a.X;
        off[4 * s + i] = out_xoff(a.om, x < a.X ? x : 0) + oy;
        vv[4 * s + i] = (double)acc[4 * hh + s][Q][i];
      }
    if (a.om.of32)
      out_add_batch<float, 16>((float*)a.om.out, off, ok, a.alpha, vv);
    else
      out_add_batch<double, 16>((double*)a.om.out, off, ok, a.alpha, vv);
    // one batch at a time: interleaving batches would need registers for all
    // their addresses and values at once
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int NQ, size_t... Qs>
__device__ inline void bf16_epilogue(std::index_sequence<Qs...>, const B16Args& a, const f4 (&acc)[8][NQ], int64_t x0,
                                     int64_t y0, int wx, int wy, int lr, int lk, int sp) {
  (bf16_epilogue_q<(int)Qs, NQ>(a, acc, x0, y0, wx, wy, lr, lk, sp, NQ * 32), ...);
}

template <int NPASS, int BYB, int S, bool HILO>
__global__ __launch_bounds__(256, 1) void k_ggemm_bf16(B16Args a) {
  constexpr int NQ = BYB / 32;
  constexpr int PER = KG * (BYB / 64) / 4;          // B chunks per wave per stage
  constexpr int LW = 8 + 2 * PER;                   // DMA instructions per wave per stage
  constexpr int LWP = NPASS > 1 ? LW : LW - PER;    // the lo planes of B are skipped with one pass
  __shared__ StageMem<BYB, HILO> st[S];
  __shared__ RowTab rtab;
  static_assert(S == 2 || S == 3, "2 or 3 stages");
  int64_t bx, yb;
  int sp;
  {
    const int64_t nb = a.nbx * a.nyb * a.nsplit;
    int64_t id = blockIdx.x;
    const int64_t per = nb / 8;
    if (per * 8 == nb) id = (id % 8) * per + id / 8;
    sp = (int)(id % a.nsplit);
    id /= a.nsplit;
    if (a.order == 0) {
      yb = id / a.nbx;
      bx = id % a.nbx;
    } else {
      bx = id / a.nyb;
      yb = id % a.nyb;
    }
  }
  const int64_t sub = a.TY / BYB;
  const int64_t t = a.tile_order[yb / sub];
  const int64_t bcol = (yb % sub) * BYB;  // first column of this block within the tile
  const int64_t y0 = t * a.TY + bcol;
  uint32_t g_beg = a.tile_off[t], g_end = a.tile_off[t + 1];
  if (a.nsplit > 1) {
    const uint32_t nst = (g_end - g_beg) / KR;
    const uint32_t per = (nst + a.nsplit - 1) / a.nsplit;
    const uint32_t s0 = min(nst, per * (uint32_t)sp), s1 = min(nst, per * (uint32_t)(sp + 1));
    g_end = g_beg + s1 * KR;
    g_beg = g_beg + s0 * KR;
  }
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int wx = w & 1, wy = w >> 1;
  const int64_t x0 = bx * BX;
  StageSrc ps;
  ps.A = a.A + (bx * a.R) * AGeom<HILO>::REC + 8 * lane;
  ps.Bhi = a.Bhi;
  ps.Blo = a.Blo;
  ps.glist = (cu32*)a.glist;
  ps.TY = a.TY;
  f4 acc[8][NQ];
#pragma unroll
  for (int s = 0; s < 8; s++)
#pragma unroll
    for (int q = 0; q < NQ; q++) acc[s][q] = f4{0, 0, 0, 0};
  // Pipeline: stage i lives in buffer i % S.  Each step waits for this
  // wave's DMA of the current stage (the next one may stay in flight),
  // barriers (every wave's DMA of the stage landed; every wave is done with
  // the buffer computed last step), refills that buffer with the stage S-1
  // ahead, then runs the MFMAs of the current stage.  The fences around the
  // barrier order LDS accesses only, so the explicit vmcnt wait is the only
  // wait on the DMA.
  // Vector-memory issue order per iteration: the row-index DMA of the stage
  // S ahead (clamped to a stage of this workgroup, so always issued), then
  // the DMAs of the stage S-1 ahead.  The wait at the top of an iteration
  // (all but the last LWP operations, the next stage's DMAs) so also
  // completes the index DMA issued just before them, one iteration earlier;
  // each wave reads its own indices, so no barrier is needed before the read.
  uint32_t k = g_beg;
  if (k < g_end) {
#pragma unroll
    for (int s = 0; s < S - 1; s++) {
      if (s == S - 2) row_dma(ps, k + (S - 1) * KR < g_end ? k + (S - 1) * KR : k, w, lane, rtab, (S - 1) & 1);
      if (k + s * KR < g_end)
        stage_issue<NPASS, BYB, HILO>(ps, k + s * KR, bcol, st[s], w, lane, row_scalar(ps, k + s * KR, w));
    }
    int cur = 0, slot = (S - 1) & 1;  // slot of the indices of stage (k - g_beg)/KR + S - 1
    for (; k < g_end; k += KR) {
      if (S > 2 && k + KR < g_end)
        wait_vm<(S > 2 ? LWP : 0)>();
      else
        wait_vm<0>();
      // the DMA'd indices are read only after the wait above
      asm volatile("" ::: "memory");
      const RowIdx ri = row_lds(rtab, slot, w);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
      const int nxt = cur == 0 ? S - 1 : cur - 1;  // (cur + S - 1) % S
      if (k + (S - 1) * KR < g_end) {
        row_dma(ps, k + S * KR < g_end ? k + S * KR : k, w, lane, rtab, slot ^ 1);
        stage_issue<NPASS, BYB, HILO>(ps, k + (S - 1) * KR, bcol, st[nxt], w, lane, ri);
      }
      stage_compute<NPASS, BYB, HILO>(st[cur], acc, wx, wy, lr, lk);
      cur = cur == S - 1 ? 0 : cur + 1;
      slot ^= 1;
    }
  }
  // D layout: lane holds column lane%16 of each 16x16 block, rows 4*(lane/16)+i
  bf16_epilogue<NQ>(std::make_index_sequence<NQ>{}, a, acc, x0, y0, wx, wy, lr, lk, sp);
}

// ---------------------------------------------------------------------------
// The 256 x 128 block kernel with its LDS reads software-pipelined.
//
// Same operands, staging, block mapping and epilogue as k_ggemm_bf16 with
// BYB = 128 and 3 LDS stages; what changes is when a wave reads LDS.  A wave
// holds the MFMA fragments of a whole K step in registers (A: 8 blocks of 16
// x, B: 4 blocks of 16 y, hi and lo) and reads the next step's fragments
// while its MFMAs run on the current ones, so no MFMA waits on an LDS read
// issued after the barrier: A block s of the next step is read into the
// registers of A block s as soon as this step's MFMAs of block s are issued,
// and B fragment q (read by every block) into its registers once the last
// block's MFMAs with q are issued.  One register set, read in place: a second
// set swapped every step would need the loop unrolled by two, and the
// compiler then splits the accumulators between register files.  The
// global->LDS DMAs of the stage S steps ahead are issued in pieces between
// the MFMA blocks (sched_barrier keeps the pieces where they are); one
// barrier per K step remains.
template <int NPASS>
struct PA {
  bf8 ah[8];
  bf8 al[NPASS > 1 ? 8 : 1];
};
template <int NPASS>
struct PB {
  uint4 bh[4];
  uint4 bl[NPASS > 1 ? 4 : 1];
};

// A fragment s (x block wx*128 + 16 s) of one stage
template <int NPASS, bool HILO>
__device__ inline void frag_read_a(const StageMem<128, HILO>& sm, PA<NPASS>& f, int s, uint32_t r0, uint32_t r1,
                                   int wx) {
  const uint32_t c = (uint32_t)((wx * 128 + 16 * s) * 2);
  f.ah[s] = cat_bf8(tr_read(sm, r0 + c), tr_read(sm, r1 + c));
  if (NPASS > 1) f.al[s] = cat_bf8(tr_read(sm, r0 + BX * 2 + c), tr_read(sm, r1 + BX * 2 + c));
}

// B fragment q (y block wy*64 + 16 q) of one stage
template <int NPASS, bool HILO>
__device__ inline void frag_read_b(const StageMem<128, HILO>& sm, PB<NPASS>& f, int q, int wy, int lr, int lk) {
  f.bh[q] = sm.B[0][lk][wy * 64 + 16 * q + lr];
  if (NPASS > 1) f.bl[q] = sm.B[1][lk][wy * 64 + 16 * q + lr];
}

template <int NPASS>
__device__ inline void frag_mfma(const PA<NPASS>& fa, const PB<NPASS>& fb, f4 (&acc)[8][4], int s, int q) {
  acc[s][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa.ah[s], as_bf8(fb.bh[q]), acc[s][q], 0, 0, 0);
  if (NPASS > 1) {
    acc[s][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa.ah[s], as_bf8(fb.bl[q]), acc[s][q], 0, 0, 0);
    acc[s][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa.al[s], as_bf8(fb.bh[q]), acc[s][q], 0, 0, 0);
  }
}

// DMA piece s of the stage at K step k (wave w): A row 8w + s, and for s < 2
// the B chunk s (hi, and lo with three passes) -- stage_issue in pieces
template <int NPASS, bool HILO>
__device__ inline void dma_piece(const StageSrc& p, uint32_t k, int64_t bcol, StageMem<128, HILO>& sm, int w,
                                 int lane, uint32_t row, int s) {
  typedef AGeom<HILO> G;
  if (HILO || lane < 32) dma16(p.A + (int64_t)row * G::REC, lds_addr(&sm.A[0]) + G::off(8 * w + s));
  if (s < 2) {
    const int c = w * 2 + s, gl = c / 2, yc = c % 2;
    const int64_t off = (int64_t)((k >> 3) + gl) * p.TY + bcol + yc * 64 + lane;
    const uint32_t l = lds_addr(&sm.B[0][0][0]) + (uint32_t)((gl * 128 + yc * 64) * 16);
    dma16(p.Bhi + off, l);
    if (NPASS > 1) dma16(p.Blo + off, l + (uint32_t)(KG * 128 * 16));
  }
}

__device__ inline uint32_t row_of(const RowIdx& r, int s) {
  switch (s) {
    case 0: return r.lo.x;
    case 1: return r.lo.y;
    case 2: return r.lo.z;
    case 3: return r.lo.w;
    case 4: return r.hi.x;
    case 5: return r.hi.y;
    case 6: return r.hi.z;
    default: return r.hi.w;
  }
}

// Epilogue through LDS for om.vec != 0: the workgroup's 256 x 128 block of
// accumulators is staged in the stage buffers (idle once the K loop is done),
// then read back along the output's unit-stride axis, so each lane reads and
// writes 4 consecutive outputs with 16-byte (FP32) or 2 x 16-byte (FP64)
// accesses and each wave covers whole runs (512 B of FP32 / 1 KB of FP64 along
// y, 1 KB / 2 KB along x).  Same arithmetic as out_add_batch: out + (TO)(alpha
// * acc).  E rows are padded by 4 floats: the writes from the MFMA layout
// (lane groups 4 rows apart) stay off one bank, the 16-byte reads aligned.
constexpr int EPY = 128 + 4;  // pitch (floats) of E[x][y] for y runs
constexpr int EPX = 256 + 4;  // pitch of E[y][x] for x runs

template <typename TO>
__device__ inline void rmw4(TO* __restrict__ out, int64_t off, double alpha, const float4& v) {
  if constexpr (sizeof(TO) == 4) {
    float4* p = (float4*)(out + off);
    float4 c = *p;
    c.x = c.x + (float)(alpha * (double)v.x);
    c.y = c.y + (float)(alpha * (double)v.y);
    c.z = c.z + (float)(alpha * (double)v.z);
    c.w = c.w + (float)(alpha * (double)v.w);
    *p = c;
  } else {
    double2* p = (double2*)(out + off);
    double2 c0 = p[0], c1 = p[1];
    c0.x = c0.x + alpha * (double)v.x;
    c0.y = c0.y + alpha * (double)v.y;
    c1.x = c1.x + alpha * (double)v.z;
    c1.y = c1.y + alpha * (double)v.w;
    p[0] = c0;
    p[1] = c1;
  }
}

template <typename TO>
__device__ inline void rmw1(TO* __restrict__ out, int64_t off, double alpha, float v) {
  out[off] = out[off] + (TO)(alpha * (double)v);
}

__device__ inline int64_t out_yoff(const OutDev& om, int64_t y) {
  const uint32_t yy = (uint32_t)y;
  return (int64_t)(yy / om.Yl) * om.oyh + (int64_t)(yy % om.Yl) * om.oyl;
}

template <typename TO>
__device__ inline void staged_store(const B16Args& a, const float* E, int64_t x0, int64_t y0, int tid) {
  TO* out = (TO*)a.om.out;
  if (a.om.vec == 1) {
    // row r of x, 32 chunks of 4 y; a wave covers two rows
    const int c = tid & 31;
    for (int r = tid >> 5; r < 256; r += 8) {
      const int64_t x = x0 + r;
      if (x >= a.X) break;
      const int64_t xo = out_xoff(a.om, x), y = y0 + 4 * c;
      const float4 v = *(const float4*)&E[r * EPY + 4 * c];
      if (y + 3 < a.Y) {
        rmw4<TO>(out, xo + out_yoff(a.om, y), a.alpha, v);
      } else {
        const float vv[4] = {v.x, v.y, v.z, v.w};
        for (int j = 0; j < 4; j++)
          if (y + j < a.Y) rmw1<TO>(out, xo + out_yoff(a.om, y + j), a.alpha, vv[j]);
      }
    }
  } else {
    // column y, 64 chunks of 4 x; a wave covers one column
    const int c = tid & 63;
    for (int yl = tid >> 6; yl < 128; yl += 4) {
      const int64_t y = y0 + yl;
      if (y >= a.Y) break;
      const int64_t yo = out_yoff(a.om, y), x = x0 + 4 * c;
      const float4 v = *(const float4*)&E[yl * EPX + 4 * c];
      if (x + 3 < a.X) {
        rmw4<TO>(out, out_xoff(a.om, x) + yo, a.alpha, v);
      } else {
        const float vv[4] = {v.x, v.y, v.z, v.w};
        for (int j = 0; j < 4; j++)
          if (x + j < a.X) rmw1<TO>(out, out_xoff(a.om, x + j) + yo, a.alpha, vv[j]);
      }
    }
  }
}

// The accumulators of wave (wx, wy) into E (layout by om.vec), then the
// staged read-modify-write.  The caller has finished every LDS read of the
// stage buffers E aliases and has no DMA in flight.
__device__ inline void bf16_epilogue_staged(const B16Args& a, const f4 (&acc)[8][4], float* E, int64_t x0,
                                            int64_t y0, int wx, int wy, int lr, int lk, int tid) {
  __syncthreads();
  const bool ymode = a.om.vec == 1;
#pragma unroll
  for (int s = 0; s < 8; s++)
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int xl = wx * 128 + 16 * s + 4 * lk + i, yl = wy * 64 + 16 * q + lr;
        E[ymode ? xl * EPY + yl : yl * EPX + xl] = acc[s][q][i];
      }
  __syncthreads();
  if (a.om.of32)
    staged_store<float>(a, E, x0, y0, tid);
  else
    staged_store<double>(a, E, x0, y0, tid);
}

template <int NPASS, bool HILO>
__global__ __launch_bounds__(256, 1) void k_ggemm_bf16_pl(B16Args a) {
  constexpr int S = 3;
  constexpr int LW = 8 + 2 * 2;                  // DMA instructions per wave per stage
  constexpr int LWP = NPASS > 1 ? LW : LW - 2;   // the lo planes of B are skipped with one pass
  __shared__ StageMem<128, HILO> st[S];
  __shared__ RowTab rtab;
  int64_t bx, yb;
  int sp;
  {
    const int64_t nb = a.nbx * a.nyb * a.nsplit;
    int64_t id = blockIdx.x;
    const int64_t per = nb / 8;
    if (per * 8 == nb) id = (id % 8) * per + id / 8;
    sp = (int)(id % a.nsplit);
    id /= a.nsplit;
    if (a.order == 0) {
      yb = id / a.nbx;
      bx = id % a.nbx;
    } else {
      bx = id / a.nyb;
      yb = id % a.nyb;
    }
  }
  const int64_t sub = a.TY / 128;
  const int64_t t = a.tile_order[yb / sub];
  const int64_t bcol = (yb % sub) * 128;  // first column of this block within the tile
  const int64_t y0 = t * a.TY + bcol;
  uint32_t g_beg = a.tile_off[t], g_end = a.tile_off[t + 1];
  if (a.nsplit > 1) {
    const uint32_t nst = (g_end - g_beg) / KR;
    const uint32_t per = (nst + a.nsplit - 1) / a.nsplit;
    const uint32_t s0 = min(nst, per * (uint32_t)sp), s1 = min(nst, per * (uint32_t)(sp + 1));
    g_end = g_beg + s1 * KR;
    g_beg = g_beg + s0 * KR;
  }
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int wx = w & 1, wy = w >> 1;
  const int64_t x0 = bx * BX;
  StageSrc ps;
  ps.A = a.A + (bx * a.R) * AGeom<HILO>::REC + 8 * lane;
  ps.Bhi = a.Bhi;
  ps.Blo = a.Blo;
  ps.glist = (cu32*)a.glist;
  ps.TY = a.TY;
  typedef AGeom<HILO> G;
  // this lane's transposed-read addresses (rows 8*lk + lr/4 and +4)
  const uint32_t r0 = G::off(8 * lk + (lr >> 2)) + (uint32_t)((lr & 3) * 8);
  const uint32_t r1 = G::off(8 * lk + 4 + (lr >> 2)) + (uint32_t)((lr & 3) * 8);
  f4 acc[8][4];
#pragma unroll
  for (int s = 0; s < 8; s++)
#pragma unroll
    for (int q = 0; q < 4; q++) acc[s][q] = f4{0, 0, 0, 0};
  if (g_beg < g_end) {
    // Prologue: stages 0..S-1 (scalar row indices), the row-index DMA of
    // stage S just before the DMAs of stage S-1; then stage 0's fragments.
    const uint32_t k0 = g_beg;
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t ks = k0 + s * KR;
      if (s == S - 1 && ks < g_end) row_dma(ps, k0 + S * KR < g_end ? k0 + S * KR : k0, w, lane, rtab, S & 1);
      if (ks < g_end) {
        const RowIdx ri = row_scalar(ps, ks, w);
#pragma unroll
        for (int i = 0; i < 8; i++) dma_piece<NPASS, HILO>(ps, ks, bcol, st[s], w, lane, row_of(ri, i), i);
      }
    }
    if (k0 + 2 * KR < g_end)
      wait_vm<2 * LWP + 1>();
    else if (k0 + KR < g_end)
      wait_vm<LWP>();
    else
      wait_vm<0>();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    PA<NPASS> fa;
    PB<NPASS> fb;
#pragma unroll
    for (int s = 0; s < 8; s++) frag_read_a<NPASS, HILO>(st[0], fa, s, r0, r1, wx);
#pragma unroll
    for (int q = 0; q < 4; q++) frag_read_b<NPASS, HILO>(st[0], fb, q, wy, lr, lk);

    // Step j (K rows [k, k+32)): its fragments are in fa and fb.  Vector-memory
    // order per step: the row-index DMA of step j+S+1, then the LWP DMAs of
    // step j+S.  The wait at the top (all but the last LWP operations) so
    // completes step j+1's DMAs, issued two steps earlier, and the indices
    // of step j+S, issued first in the previous step.
    uint32_t k = k0;
    int cur = 0, slot = S & 1;
    while (k < g_end) {
      const bool more2 = k + 2 * KR < g_end, issue = k + S * KR < g_end;
      if (more2)
        wait_vm<LWP>();
      else
        wait_vm<0>();
      asm volatile("" ::: "memory");  // the DMA'd indices are read after the wait
      const RowIdx ri = row_lds(rtab, slot, w);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
      const int nb = cur == S - 1 ? 0 : cur + 1;  // buffer of step j+1
      const uint32_t kd = k + S * KR;             // step whose DMAs go into buffer cur
#pragma unroll
      for (int s = 0; s < 8; s++) {
        if (issue) {
          if (s == 0) row_dma(ps, kd + KR < g_end ? kd + KR : k, w, lane, rtab, slot ^ 1);
          dma_piece<NPASS, HILO>(ps, kd, bcol, st[cur], w, lane, row_of(ri, s), s);
        }
        // next step's fragments, read unconditionally: past the last step
        // buffer nb holds stale data whose fragments are never used
#pragma unroll
        for (int q = 0; q < 4; q++) {
          frag_mfma<NPASS>(fa, fb, acc, s, q);
          if (s == 7) frag_read_b<NPASS, HILO>(st[nb], fb, q, wy, lr, lk);
        }
        frag_read_a<NPASS, HILO>(st[nb], fa, s, r0, r1, wx);
        __builtin_amdgcn_sched_barrier(0);
      }
      cur = nb;
      slot ^= 1;
      k += KR;
    }
  }
  // the staged epilogue needs the hi|lo stages' 148 KB (hi-only stages hold 100 KB)
  if constexpr (HILO) {
    if (a.nsplit == 1 && a.om.vec != 0) {
      static_assert(sizeof(st) >= sizeof(float) * 256 * EPY && sizeof(st) >= sizeof(float) * 128 * EPX,
                    "the staged epilogue fits in the stage buffers");
      bf16_epilogue_staged(a, acc, reinterpret_cast<float*>(&st[0]), x0, y0, wx, wy, lr, lk, tid);
      return;
    }
  }
  bf16_epilogue<4>(std::make_index_sequence<4>{}, a, acc, x0, y0, wx, wy, lr, lk, sp);
}

// out[x, y] += alpha * sum_s partial[s][x][y], splits added in order
__global__ void k_splitk_reduce(const float* __restrict__ partial, int nsplit, int64_t X, int64_t Y, OutDev om,
                                double alpha) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= X * Y) return;
  float acc = 0.0f;
  for (int s = 0; s < nsplit; s++) acc += partial[(int64_t)s * X * Y + e];
  const uint32_t xx = (uint32_t)(e / Y), yy = (uint32_t)(e % Y);
  out_add(om,
          (int64_t)(xx / om.Xl) * om.oxh + (int64_t)(xx % om.Xl) * om.oxl + (int64_t)(yy / om.Yl) * om.oyh +
              (int64_t)(yy % om.Yl) * om.oyl,
          alpha, (double)acc);
}

// ---------------------------------------------------------------------------
// FP64 / FP32 gathered contraction through LDS (v_mfma_f64_16x16x4f64,
// v_mfma_f32_16x16x4f32).  A workgroup owns BXF x 64 outputs: x in
// [bx*BXF, +BXF) with BXF = 1 KB of A per row (128 doubles or 256 floats) and
// one 64-column view tile.  Per stage it gathers 16 rows: each row's 1 KB
// segment of A is one global->LDS DMA (rows at a padded pitch so the four
// rows one MFMA K step reads sit on different banks), the tile's 16 B rows
// are contiguous bytes of the view (4 or 8 KB).  Three stages in flight; the
// waves (2 x 2) own BXF/2 x 32 outputs each.  Each A element of a gathered
// row is fetched once per workgroup (the register-loading kernel above read
// it once per lane group).  Needs X a multiple of 16 bytes' worth of
// elements and a 16-byte aligned A, so every 16-byte chunk lies inside one
// row.
template <typename T>
struct FpGeom {
  static constexpr int BXF = 1024 / (int)sizeof(T);  // x per workgroup: one 1 KB segment per row
  static constexpr int KR = 16;                       // gathered rows per stage (4 MFMA K steps)
  static constexpr int APITCH = 1024 + (sizeof(T) == 8 ? 128 : 64);
  static constexpr int BROW = GT_Y * (int)sizeof(T);  // bytes of one view row
  static constexpr int ABYTES = KR * APITCH;
  static constexpr int BBYTES = KR * BROW;
  static constexpr int PERB = BBYTES / 1024 / 4;      // view DMAs per wave per stage
  static constexpr int WX = BXF / 2, WY = GT_Y / 2;   // outputs per wave
  static constexpr int MS = WX / 16, NQ = WY / 16;    // MFMA blocks per wave
};

template <typename T>
struct FpStage {
  uint8_t A[FpGeom<T>::ABYTES];
  uint8_t B[FpGeom<T>::BBYTES];
};

struct FArgs {
  const void* A;  // [R][X]
  int64_t X, Y;
  const uint32_t* tile_off;
  const uint32_t* tile_order;  // tiles by decreasing row count
  const uint32_t* rowlist;
  const void* B;  // view [rows][64]
  int64_t nbx, ntiles;
  OutDev om;
  double alpha;
  int order;
};

template <typename T>
__device__ inline T lds_ld(const uint8_t* base, uint32_t off) {
  return *(const T*)(base + off);
}

// Row indices of the FP pipeline: wave w stages rows 4w..4w+3 of a stage
// (16 bytes).  As in the BF16 kernel, inside the loop they reach the wave
// through a two-slot LDS table, DMA'd by lane 0 one iteration before they are
// used (in vmcnt order with the stage DMAs); the prologue reads them through
// the scalar cache.
__device__ inline uint4 fp_rows_scalar(cu32* rl, uint32_t k, int w) {
  return uint4{rl[k + 4 * w], rl[k + 4 * w + 1], rl[k + 4 * w + 2], rl[k + 4 * w + 3]};
}

__device__ inline void fp_rows_dma(cu32* rl, uint32_t k, int w, int lane, uint4 (&tab)[2][4], int slot) {
  if (lane == 0) dma16((const uint32_t*)rl + k + 4 * w, lds_addr(&tab[slot][w]));
}

template <typename T>
__device__ inline void fp_issue(const FArgs& a, uint4 rows, int64_t aoff, uint32_t k, FpStage<T>& sm, int w,
                                int lane) {
  typedef FpGeom<T> G;
  const uint32_t la = lds_addr(&sm.A[0]), lb = lds_addr(&sm.B[0]);
  const T* Ab = (const T*)a.A;
  const uint32_t rr[4] = {rows.x, rows.y, rows.z, rows.w};
#pragma unroll
  for (int i = 0; i < 4; i++) dma16(Ab + (int64_t)rr[i] * a.X + aoff, la + (uint32_t)((4 * w + i) * G::APITCH));
  const uint8_t* Bb = (const uint8_t*)a.B + (int64_t)k * G::BROW;
#pragma unroll
  for (int j = 0; j < G::PERB; j++) {
    const int c = w * G::PERB + j;
    dma16(Bb + c * 1024 + 16 * lane, lb + (uint32_t)(c * 1024));
  }
}

template <typename T>
__device__ inline void fp_compute(const FpStage<T>& sm, typename Acc<T>::type (&acc)[FpGeom<T>::MS][FpGeom<T>::NQ],
                                  int wx, int wy, int lr, int lk) {
  typedef FpGeom<T> G;
#pragma unroll
  for (int kk = 0; kk < G::KR / 4; kk++) {
    T av[G::MS], bv[G::NQ];
    const uint32_t ar = (uint32_t)((4 * kk + lk) * G::APITCH), br = (uint32_t)((4 * kk + lk) * G::BROW);
#pragma unroll
    for (int s = 0; s < G::MS; s++) av[s] = lds_ld<T>(sm.A, ar + (uint32_t)((wx * G::WX + 16 * s + lr) * sizeof(T)));
#pragma unroll
    for (int q = 0; q < G::NQ; q++) bv[q] = lds_ld<T>(sm.B, br + (uint32_t)((wy * G::WY + 16 * q + lr) * sizeof(T)));
#pragma unroll
    for (int s = 0; s < G::MS; s++)
#pragma unroll
      for (int q = 0; q < G::NQ; q++) acc[s][q] = Acc<T>::mfma(av[s], bv[q], acc[s][q]);
  }
}

// FP64 epilogue through LDS (om.vec != 0, FP64 output): the 128 x 64 block of
// accumulators is staged in the stage buffers and read back along the
// output's unit-stride axis, 2 consecutive outputs (16 bytes) per lane and
// whole runs per wave (512 B along y, 1 KB along x); same arithmetic as
// out_add_batch.  Rows padded by 2 doubles (16-byte aligned, off one bank).
constexpr int FEPY = 64 + 2;   // pitch (doubles) of E[x][y] for y runs
constexpr int FEPX = 128 + 2;  // pitch of E[y][x] for x runs

__device__ inline void rmw2(double* __restrict__ out, int64_t off, double alpha, const double2& v) {
  double2* p = (double2*)(out + off);
  double2 c = *p;
  c.x = c.x + alpha * v.x;
  c.y = c.y + alpha * v.y;
  *p = c;
}

__device__ inline void fp64_epilogue_staged(const FArgs& a, const d4 (&acc)[4][2], double* E, int64_t x0, int64_t y0,
                                            int wx, int wy, int lr, int lk, int tid) {
  __syncthreads();
  const bool ymode = a.om.vec == 1;
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int xl = wx * 64 + 16 * s + Acc<double>::row(lk, i), yl = wy * 32 + 16 * q + lr;
        E[ymode ? xl * FEPY + yl : yl * FEPX + xl] = acc[s][q][i];
      }
  __syncthreads();
  double* out = (double*)a.om.out;
  if (ymode) {
    // row x, 32 chunks of 2 y; a wave covers two rows
    const int c = tid & 31;
    for (int r = tid >> 5; r < 128; r += 8) {
      const int64_t x = x0 + r;
      if (x >= a.X) break;
      const int64_t xo = out_xoff(a.om, x), y = y0 + 2 * c;
      const double2 v = *(const double2*)&E[r * FEPY + 2 * c];
      if (y + 1 < a.Y) {
        rmw2(out, xo + out_yoff(a.om, y), a.alpha, v);
      } else if (y < a.Y) {
        const int64_t off = xo + out_yoff(a.om, y);
        out[off] = out[off] + a.alpha * v.x;
      }
    }
  } else {
    // column y, 64 chunks of 2 x; a wave covers one column
    const int c = tid & 63;
    for (int yl = tid >> 6; yl < 64; yl += 4) {
      const int64_t y = y0 + yl;
      if (y >= a.Y) break;
      const int64_t yo = out_yoff(a.om, y), x = x0 + 2 * c;
      const double2 v = *(const double2*)&E[yl * FEPX + 2 * c];
      if (x + 1 < a.X) {
        rmw2(out, out_xoff(a.om, x) + yo, a.alpha, v);
      } else if (x < a.X) {
        const int64_t off = out_xoff(a.om, x) + yo;
        out[off] = out[off] + a.alpha * v.x;
      }
    }
  }
}

template <typename T, int S>
__global__ __launch_bounds__(256, 2) void k_ggemm_fp(FArgs a) {
  typedef FpGeom<T> G;
  typedef typename Acc<T>::type AT;
  constexpr int LW = 4 + G::PERB;  // DMA instructions per wave per stage
  static_assert(S == 3, "the row-index prefetch assumes three stages");
  __shared__ FpStage<T> st[S];
  __shared__ uint4 rtab[2][4];
  int64_t bx, ti;
  xcd_remap(a.nbx, a.ntiles, a.order, bx, ti);
  const int64_t t = a.tile_order[ti];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int wx = w & 1, wy = w >> 1;
  const uint32_t k_beg = a.tile_off[t], k_end = a.tile_off[t + 1];
  const int64_t x0 = bx * G::BXF;
  // this lane's 16-byte chunk of each row segment; chunks past X read the
  // row's start instead (their columns are never stored)
  const int64_t xl = x0 + lane * (16 / (int64_t)sizeof(T));
  const int64_t aoff = xl < a.X ? xl : 0;
  cu32* rl = (cu32*)a.rowlist;
  AT acc[G::MS][G::NQ];
#pragma unroll
  for (int s = 0; s < G::MS; s++)
#pragma unroll
    for (int q = 0; q < G::NQ; q++) acc[s][q] = AT{0, 0, 0, 0};
  // Vector-memory order per iteration: the row-index DMA of the stage S
  // ahead (clamped to a stage of this tile, so always issued), then the LW
  // DMAs of the stage S-1 ahead; the wait at the top (all but the last LW
  // operations) so also completes the indices DMA'd one iteration earlier.
  uint32_t k = k_beg;
  if (k < k_end) {
#pragma unroll
    for (int s = 0; s < S - 1; s++) {
      if (s == S - 2) fp_rows_dma(rl, k + (S - 1) * G::KR < k_end ? k + (S - 1) * G::KR : k, w, lane, rtab, (S - 1) & 1);
      if (k + s * G::KR < k_end) fp_issue<T>(a, fp_rows_scalar(rl, k + s * G::KR, w), aoff, k + s * G::KR, st[s], w, lane);
    }
    int cur = 0, slot = (S - 1) & 1;  // slot of the indices of stage (k - k_beg)/KR + S - 1
    for (; k < k_end; k += G::KR) {
      if (k + G::KR < k_end)
        wait_vm<LW * (S - 2)>();
      else
        wait_vm<0>();
      asm volatile("" ::: "memory");  // the DMA'd indices are read after the wait
      const uint4 rows = rtab[slot][w];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
      const int nxt = cur == 0 ? S - 1 : cur - 1;
      if (k + (S - 1) * G::KR < k_end) {
        fp_rows_dma(rl, k + S * G::KR < k_end ? k + S * G::KR : k, w, lane, rtab, slot ^ 1);
        fp_issue<T>(a, rows, aoff, k + (S - 1) * G::KR, st[nxt], w, lane);
      }
      fp_compute<T>(st[cur], acc, wx, wy, lr, lk);
      cur = cur == S - 1 ? 0 : cur + 1;
      slot ^= 1;
    }
  }
  if constexpr (sizeof(T) == 8) {
    if (a.om.vec != 0 && !a.om.of32) {
      static_assert(sizeof(st) >= sizeof(double) * 128 * FEPY && sizeof(st) >= sizeof(double) * 64 * FEPX,
                    "the staged epilogue fits in the stage buffers");
      static_assert(G::MS == 4 && G::NQ == 2, "FP64 wave tile of 64 x 32");
      fp64_epilogue_staged(a, acc, reinterpret_cast<double*>(&st[0]), x0, t * GT_Y, wx, wy, lr, lk, tid);
      return;
    }
  }
#pragma unroll
  for (int q = 0; q < G::NQ; q++) {
    const int64_t y = t * GT_Y + wy * G::WY + 16 * q + lr;
    if (y >= a.Y) continue;
    const uint32_t yy = (uint32_t)y;
    const int64_t oy = (int64_t)(yy / a.om.Yl) * a.om.oyh + (int64_t)(yy % a.om.Yl) * a.om.oyl;
#pragma unroll
    for (int h = 0; h < G::MS / 4; h++) {
      int64_t off[16];
      bool ok[16];
      double vv[16];
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int64_t x = x0 + wx * G::WX + 16 * (4 * h + s) + Acc<T>::row(lk, i);
          ok[4 * s + i] = x < a.X;
          off[4 * s + i] = out_xoff(a.om, x < a.X ? x : 0) + oy;
          vv[4 * s + i] = (double)acc[4 * h + s][q][i];
        }
      if (a.om.of32)
        out_add_batch<float, 16>((float*)a.om.out, off, ok, a.alpha, vv);
      else
        out_add_batch<double, 16>((double*)a.om.out, off, ok, a.alpha, vv);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

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
  {
    // staged FP64 epilogue: 2 consecutive y (or x) inside one level, unit
    // stride, every other stride even and the base 16-byte aligned
    const bool al = ((uintptr_t)m.out % 16) == 0 && m.oxh % 2 == 0 && m.oyh % 2 == 0;
    if (al && m.oyl == 1 && v->Yl % 2 == 0 && m.oxl % 2 == 0)
      om.vec = 1;
    else if (al && m.oxl == 1 && m.Xl % 2 == 0 && m.oyl % 2 == 0)
      om.vec = 2;
    else
      om.vec = 0;
  }
  const size_t esz = adtype == DT_F64 ? 8 : 4;
  if (g_fp_lds && v->TY == GT_Y && X % (int64_t)(16 / esz) == 0 && ((uintptr_t)A & 15) == 0 && v->tile_order) {
    FArgs a;
    a.A = A;
    a.X = X;
    a.Y = v->Y;
    a.tile_off = v->tile_off;
    a.tile_order = v->tile_order;
    a.rowlist = v->rowlist;
    a.B = v->B;
    a.ntiles = v->ntiles;
    a.om = om;
    a.alpha = alpha;
    a.order = g_order;
    if (adtype == DT_F64) {
      a.nbx = ceil_div(X, FpGeom<double>::BXF);
      k_ggemm_fp<double, 3><<<a.nbx * a.ntiles, 256, 0, s>>>(a);
    } else {
      a.nbx = ceil_div(X, FpGeom<float>::BXF);
      k_ggemm_fp<float, 3><<<a.nbx * a.ntiles, 256, 0, s>>>(a);
    }
    FRICC_HIP(hipGetLastError());
    return OK;
  }
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

int g_bf16_width = 0;  // 0: by operand size

int bf16_view_width(int64_t R, int64_t X) {
  if (g_bf16_width > 0) return g_bf16_width;
  // operands too large to stay in the caches are streamed once per tile:
  // wider tiles cut that traffic (A is R*X*4 bytes) but hold more active rows
  // each (lower fill, more MFMA work per entry).  Measured at o=32, v=256,
  // m=4e6 after ten iterations: 128 columns for every operand -- the O^3V^3
  // ones (0.27 GB), Gx/Gy (2.1 GB; 0.84 against 0.91 ms at 256) and the
  // ladder (17 GB; 5.9 against 6.4 ms; 256 was ahead only over the first
  // iterations, before the samples spread over more rows).  One width also
  // lets the ladder, Gx and Gy share one view of T.
  (void)R;
  (void)X;
  return 128;
}

int gathered_gemm_bf16(const View* v, const uint16_t* A2, int64_t R, int64_t X, const OutMap& m, double alpha,
                       int npass, hipStream_t s, int hilo) {
  FRICC_CHECK((v->dtype == DT_BF16 || v->dtype == DT_BF16X1) && v->TY % 128 == 0,
              "gathered_gemm_bf16: needs a BF16 view with tile width k*128");
  FRICC_CHECK(X > 0 && X < ((int64_t)1 << 31), "gathered_gemm_bf16: bad X");
  FRICC_CHECK(npass == 1 || npass == 3, "gathered_gemm_bf16: npass must be 1 or 3");
  FRICC_CHECK(hilo || npass == 1, "gathered_gemm_bf16: three passes need the lo plane of A");
  FRICC_CHECK(npass == 1 || v->Blo != nullptr, "gathered_gemm_bf16: three passes need the lo plane of the view");
  if (v->ntiles == 0) return OK;
  B16Args a;
  a.om.out = m.out;
  a.om.of32 = m.of32;
  a.om.Xl = (uint32_t)m.Xl;
  a.om.Yl = (uint32_t)v->Yl;
  a.om.oxh = m.oxh;
  a.om.oxl = m.oxl;
  a.om.oyh = m.oyh;
  a.om.oyl = m.oyl;
  {
    // staged epilogue: 4 consecutive y (or x) inside one level, unit stride,
    // every other stride and the base 16-byte aligned
    const bool al = ((uintptr_t)m.out % 16) == 0 && m.oxh % 4 == 0 && m.oyh % 4 == 0;
    if (al && m.oyl == 1 && v->Yl % 4 == 0 && m.oxl % 4 == 0)
      a.om.vec = 1;
    else if (al && m.oxl == 1 && m.Xl % 4 == 0 && m.oyl % 4 == 0)
      a.om.vec = 2;
    else
      a.om.vec = 0;
  }
  a.A = A2;
  a.R = R;
  a.X = X;
  a.Y = v->Y;
  a.tile_off = v->tile_off;
  a.tile_order = v->tile_order;
  a.glist = v->rowlist;
  a.Bhi = (const uint4*)v->B;
  a.Blo = (const uint4*)v->Blo;
  a.TY = v->TY;
  a.nbx = ceil_div(X, BX);
  // tiles of 256+ columns: 256 x 256 output blocks, 2 stages (131 KB of LDS);
  // else 256 x 128 blocks, 3 stages (149 KB); one workgroup per CU.  The
  // larger block reads a third fewer operand bytes per MFMA from L2.
  const int byb = v->TY % 256 == 0 ? 256 : 128;
  a.nyb = v->ntiles * (v->TY / byb);
  a.order = g_order;
  a.alpha = alpha;
  a.partial = nullptr;
  // grids of fewer workgroups than CUs: split each tile's K range, partial
  // sums added in order (at 256 workgroups, the t1-dressing contractions at
  // o=32 v=256, splitting in two was slower: 0.98 against 0.91 ms, plus the
  // reduce)
  const int64_t nb0 = a.nbx * a.nyb;
  int nsplit = 1;
  if (g_split > 0) {
    nsplit = g_split;
  } else if (nb0 < 256) {
    nsplit = (int)std::min<int64_t>(16, ceil_div(512, nb0));
    while (nsplit > 1 && (double)nsplit * X * v->Y * 4.0 > 1.0e9) nsplit--;
  }
  a.nsplit = nsplit;
  if (nsplit > 1) {
    void* p;
    FRICC_TRY(ws_get(32, (size_t)nsplit * X * v->Y * 4, &p));
    a.partial = (float*)p;
  }
  const int64_t nb = nb0 * nsplit;
  if (byb == 256) {
    if (!hilo)
      k_ggemm_bf16<1, 256, 2, false><<<nb, 256, 0, s>>>(a);
    else if (npass == 3)
      k_ggemm_bf16<3, 256, 2, true><<<nb, 256, 0, s>>>(a);
    else
      k_ggemm_bf16<1, 256, 2, true><<<nb, 256, 0, s>>>(a);
  } else {
    if (!hilo)
      k_ggemm_bf16_pl<1, false><<<nb, 256, 0, s>>>(a);
    else if (npass == 3)
      k_ggemm_bf16_pl<3, true><<<nb, 256, 0, s>>>(a);
    else
      k_ggemm_bf16_pl<1, true><<<nb, 256, 0, s>>>(a);
  }
  FRICC_HIP(hipGetLastError());
  if (nsplit > 1) k_splitk_reduce<<<ceil_div(X * v->Y, 256), 256, 0, s>>>(a.partial, nsplit, X, v->Y, a.om, alpha);
  FRICC_HIP(hipGetLastError());
  return OK;
}

// A[R][X] (f64 or f32) -> records [X/256][R][hi 256 | lo 256] bf16 (X
// rounded up to a multiple of 256, pad columns zero); hi = bf16(a),
// lo = bf16(a - hi)
// With B: the split of alpha*A + beta*B, formed in the source precision (one
// rounding of the sum of the two products; both exact for alpha, beta = 2, -1).
// hilo 0: hi-only records [X/256][R][hi 256]
// With B and out_a / out_b: the records of A and of B alone as well, from the
// same reads (hi and lo, like out).
__device__ inline void put_split(__hip_bfloat16* __restrict__ o, int64_t rec, int64_t rec_len, int64_t xi, int hilo,
                                 float v) {
  const __hip_bfloat16 h = __float2bfloat16(v);
  o[rec * rec_len + xi] = h;
  if (hilo) o[rec * rec_len + 256 + xi] = __float2bfloat16(v - __bfloat162float(h));
}

template <typename T>
__global__ void k_split_bf16(const T* __restrict__ A, const T* __restrict__ B, T alpha, T beta, int64_t R,
                             int64_t X, int64_t n, int hilo, __hip_bfloat16* __restrict__ out,
                             __hip_bfloat16* __restrict__ out_a, __hip_bfloat16* __restrict__ out_b) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t rec_len = hilo ? 512 : 256;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    // e = (xb * R + r) * 256 + xi
    const int64_t xi = e & 255, rec = e >> 8;
    const int64_t r = rec % R, x = (rec / R) * 256 + xi;
    T a = 0, b = 0;
    if (x < X) {
      a = A[r * X + x];
      if (B) b = B[r * X + x];
    }
    const float v = B ? (float)(alpha * a + beta * b) : (float)a;
    put_split(out, rec, rec_len, xi, hilo, v);
    if (out_a) put_split(out_a, rec, rec_len, xi, hilo, (float)a);
    if (out_b) put_split(out_b, rec, rec_len, xi, hilo, (float)b);
  }
}

int split_bf16(const void* A, int src_dtype, int64_t R, int64_t X, uint16_t* A2, hipStream_t s, const void* B,
               double alpha, double beta, int hilo, uint16_t* A2a, uint16_t* A2b) {
  const int64_t n = R * bf16_xblocks(X) * 256;
  if (n == 0) return OK;
  FRICC_CHECK(B || (!A2a && !A2b), "split_bf16: separate records need B");
  const int64_t grid = std::min<int64_t>(ceil_div(n, 256), 256 * 64);
  if (src_dtype == DT_F64)
    k_split_bf16<double><<<grid, 256, 0, s>>>((const double*)A, (const double*)B, alpha, beta, R, X, n, hilo,
                                              (__hip_bfloat16*)A2, (__hip_bfloat16*)A2a, (__hip_bfloat16*)A2b);
  else if (src_dtype == DT_F32)
    k_split_bf16<float><<<grid, 256, 0, s>>>((const float*)A, (const float*)B, (float)alpha, (float)beta, R, X,
                                             n, hilo, (__hip_bfloat16*)A2, (__hip_bfloat16*)A2a,
                                             (__hip_bfloat16*)A2b);
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
