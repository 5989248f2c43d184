// Device-resident SparseTensor4d storage and its gathered views.
//
// Storage replaces the AoS vector<array<size_t,4>> + vector<double> of
// include/sparse_tensor.hpp:15-18 (40 B/entry) with SoA uint32 flat index +
// FP64 value (12 B/entry) in compression order.  Instead of rescanning every
// entry per output block (src/sparse_tensor.cpp, e.g. :63-77), each
// contraction reads a view that groups the entries by output-column tile and
// stores, per tile, the dense block of the rows that hold any entry.
#include "fricc_internal.h"

namespace fricc {

constexpr int VNT = 256;

View::~View() {
  if (tile_off) (void)hipFree(tile_off);
  if (tile_order) (void)hipFree(tile_order);
  if (rowlist) (void)hipFree(rowlist);
  if (B) (void)hipFree(B);
  if (Blo) (void)hipFree(Blo);
}

Sparse::~Sparse() {
  views.clear();
  if (dsc) (void)hipFree(dsc);
  if (idx) (void)hipFree(idx);
  if (val) (void)hipFree(val);
}

int sparse_reserve(Sparse* sp, int64_t nnz) {
  if (sp->cap >= nnz && sp->idx) return OK;
  if (sp->idx) FRICC_HIP(hipFree(sp->idx));
  if (sp->val) FRICC_HIP(hipFree(sp->val));
  sp->idx = nullptr;
  sp->val = nullptr;
  FRICC_HIP(hipMalloc(&sp->idx, (size_t)std::max<int64_t>(nnz, 1) * 4));
  FRICC_HIP(hipMalloc(&sp->val, (size_t)std::max<int64_t>(nnz, 1) * 8));
  sp->cap = nnz;
  return OK;
}

struct VGeom {
  uint32_t d1, d2, d3;
  int rh, rl, yh, yl;
  uint32_t Rl, Yl, TY;
  uint32_t R;  // rows of the slot table
};

__device__ inline void view_rc(uint32_t f, const VGeom& g, uint32_t& r, uint32_t& y) {
  uint32_t i[4];
  i[3] = f % g.d3;
  f /= g.d3;
  i[2] = f % g.d2;
  f /= g.d2;
  i[1] = f % g.d1;
  i[0] = f / g.d1;
  r = i[g.rh] * g.Rl + i[g.rl];
  y = i[g.yh] * g.Yl + i[g.yl];
}

__global__ void k_mark(const uint32_t* __restrict__ idx, int64_t nnz, VGeom g, uint32_t* __restrict__ slot) {
  int64_t e = (int64_t)blockIdx.x * VNT + threadIdx.x;
  if (e >= nnz) return;
  uint32_t r, y;
  view_rc(idx[e], g, r, y);
  slot[(int64_t)(y / g.TY) * g.R + r] = 1u;
}

// Tile t's rows are split into chunks of VCH rows, one block per (chunk,
// tile), so that views with few tiles (8 for the ladder) still fill the GPU.
constexpr uint32_t VCH = 2048;

// Marked rows per (tile, chunk): cc[t * nch + c]
__global__ void k_tile_count(const uint32_t* __restrict__ slot, uint32_t R, uint32_t nch, uint32_t* __restrict__ cc) {
  __shared__ uint32_t sm[VNT / WAVE + 1];
  const uint32_t t = blockIdx.y, c = blockIdx.x;
  const uint32_t* row = slot + (int64_t)t * R;
  const uint32_t r1 = min(R, (c + 1) * VCH);
  uint32_t n = 0;
  for (uint32_t r = c * VCH + threadIdx.x; r < r1; r += VNT) n += row[r];
  uint32_t tot;
  block_excl_scan_u32<VNT>(n, &tot, sm);
  if (threadIdx.x == 0) cc[(int64_t)t * nch + c] = tot;
}

// Per tile: its rows padded to the K step
__global__ void k_tile_total(const uint32_t* __restrict__ cc, uint32_t nch, int64_t ntiles, int kstep,
                             uint32_t* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * VNT + threadIdx.x;
  if (t >= ntiles) return;
  uint32_t tot = 0;
  for (uint32_t c = 0; c < nch; c++) tot += cc[t * nch + c];
  cnt[t] = (tot + kstep - 1) / kstep * kstep;
}

// Compaction of one (tile, chunk): marked rows in increasing order, from the
// tile's offset plus the earlier chunks' counts; the tile's last chunk also
// writes the padding rows.  slot[r] becomes the row's position in the list.
__global__ void k_tile_slots(uint32_t* __restrict__ slot, uint32_t R, uint32_t nch, const uint32_t* __restrict__ cc,
                             const uint32_t* __restrict__ tile_off, uint32_t* __restrict__ rowlist) {
  __shared__ uint32_t sm[VNT / WAVE + 1];
  const uint32_t t = blockIdx.y, c = blockIdx.x;
  uint32_t* row = slot + (int64_t)t * R;
  uint32_t run = tile_off[t];
  for (uint32_t q = 0; q < c; q++) run += cc[(int64_t)t * nch + q];
  const uint32_t rbeg = c * VCH, rend = min(R, (c + 1) * VCH);
  constexpr int IT = 8;  // VNT * IT == VCH
  static_assert(VNT * IT == VCH, "one pass of the block covers a chunk");
  {
    uint32_t f[IT];
    uint32_t s = 0;
    const uint32_t r0 = rbeg + threadIdx.x * IT;
#pragma unroll
    for (int q = 0; q < IT; q++) {
      f[q] = (r0 + q < rend) ? row[r0 + q] : 0u;
      s += f[q];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan_u32<VNT>(s, &tot, sm) + run;
#pragma unroll
    for (int q = 0; q < IT; q++) {
      if (f[q]) {
        row[r0 + q] = ex;
        rowlist[ex] = r0 + q;
        ex++;
      }
    }
    run += tot;
  }
  // padding rows (zero values) point at row 0
  if (c == nch - 1)
    for (uint32_t k = run + threadIdx.x; k < tile_off[t + 1]; k += VNT) rowlist[k] = 0u;
}

// Entries with distinct flat indices land on distinct elements of the zeroed
// block and are stored; entries that may repeat an index (dup) are added.
__device__ inline bool may_repeat(const uint32_t* dup_dev, int dup_host) {
  return dup_host != 0 || (dup_dev != nullptr && *dup_dev != 0u);
}

template <typename T>
__global__ void k_fill(const uint32_t* __restrict__ idx, const double* __restrict__ val, int64_t nnz, VGeom g,
                       const uint32_t* __restrict__ slot, const uint32_t* __restrict__ dup_dev, int dup_host,
                       T* __restrict__ B) {
  int64_t e = (int64_t)blockIdx.x * VNT + threadIdx.x;
  if (e >= nnz) return;
  uint32_t r, y;
  view_rc(idx[e], g, r, y);
  uint32_t t = y / g.TY;
  uint32_t k = slot[(int64_t)t * g.R + r];
  T* dst = &B[(int64_t)k * g.TY + (y % g.TY)];
  if (may_repeat(dup_dev, dup_host))
    atomicAdd(dst, (T)val[e]);
  else
    *dst = (T)val[e];
}

// BF16 view: accumulate into an FP32 buffer laid out [k/8][TY][8] by gathered
// row k (eight consecutive gathered rows interleaved per column), split
// into the bf16 planes afterwards
__global__ void k_fill_grouped(const uint32_t* __restrict__ idx, const double* __restrict__ val, int64_t nnz, VGeom g,
                               const uint32_t* __restrict__ slot, const uint32_t* __restrict__ dup_dev, int dup_host,
                               float* __restrict__ B) {
  int64_t e = (int64_t)blockIdx.x * VNT + threadIdx.x;
  if (e >= nnz) return;
  uint32_t r, y;
  view_rc(idx[e], g, r, y);
  uint32_t t = y / g.TY;
  uint32_t k = slot[(int64_t)t * g.R + r];
  float* dst = &B[((int64_t)(k >> 3) * g.TY + (y % g.TY)) * 8 + (k & 7)];
  if (may_repeat(dup_dev, dup_host))
    atomicAdd(dst, (float)val[e]);
  else
    *dst = (float)val[e];
}

// Tiles by decreasing gathered-row count (ties by index): order[rank] = tile.
// Workgroups take tiles in this order, so the longest K ranges start first.
__global__ void k_tile_order(const uint32_t* __restrict__ tile_off, int64_t ntiles, uint32_t* __restrict__ order) {
  for (int64_t t = (int64_t)blockIdx.x * VNT + threadIdx.x; t < ntiles; t += (int64_t)gridDim.x * VNT) {
    const uint32_t c = tile_off[t + 1] - tile_off[t];
    int64_t rank = 0;
    for (int64_t u = 0; u < ntiles; u++) {
      const uint32_t cu = tile_off[u + 1] - tile_off[u];
      rank += (cu > c) || (cu == c && u < t);
    }
    order[rank] = (uint32_t)t;
  }
}

// Zero n words (a kernel rather than a memset, so a view build can be
// recorded in a graph)
__global__ void k_zero_u32(uint32_t* __restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * VNT;
  for (int64_t i = (int64_t)blockIdx.x * VNT + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// Zero the first (*rows) rows of a row-major buffer of `words` uint4 per row;
// the row count is read on the device, so the host never waits for it.
__global__ void k_zero_rows(const uint32_t* __restrict__ rows, int64_t words, uint4* __restrict__ B) {
  const int64_t n = (int64_t)(*rows) * words;
  const int64_t stride = (int64_t)gridDim.x * VNT;
  for (int64_t i = (int64_t)blockIdx.x * VNT + threadIdx.x; i < n; i += stride) B[i] = make_uint4(0, 0, 0, 0);
}

// FP32 [rows][w] -> bf16 hi = bf16(v), lo = bf16(v - hi) for the first (*rows)
// rows (row count read on the device), round to nearest even
__global__ void k_split_rows(const uint32_t* __restrict__ rows, int64_t w, const float* __restrict__ in,
                             __hip_bfloat16* __restrict__ hi, __hip_bfloat16* __restrict__ lo) {
  const int64_t n = (int64_t)(*rows) * w;
  const int64_t stride = (int64_t)gridDim.x * VNT;
  for (int64_t i = (int64_t)blockIdx.x * VNT + threadIdx.x; i < n; i += stride) {
    const float f = in[i];
    const __hip_bfloat16 h = __float2bfloat16(f);
    hi[i] = h;
    if (lo) lo[i] = __float2bfloat16(f - __bfloat162float(h));  // no lo plane for single-pass views
  }
}

// Worst-case gathered rows of a view: every tile holds at most all of its
// slot rows and at most one row per entry, each padded to the K step.
static size_t view_capacity(int64_t ntiles, int64_t slot_rows, int64_t nnz, int kstep) {
  const int64_t pad = ntiles * (kstep - 1);
  return (size_t)std::max<int64_t>(1, std::min(ntiles * slot_rows + pad, nnz + pad));
}

// Builds the view on the stream without a host round trip: storage is sized
// for the worst case (at most the dense tensor, 4 or 8 bytes per element), the
// per-tile offsets stay on the device and the GEMM reads them there.
static int build_view(Sparse* sp, View* v, hipStream_t s) {
  const int64_t* d = sp->dims;
  v->Rh = d[v->pos[0]];
  v->Rl = d[v->pos[1]];
  v->Yh = d[v->pos[2]];
  v->Yl = d[v->pos[3]];
  v->R = v->Rh * v->Rl;
  v->Y = v->Yh * v->Yl;
  v->ntiles = ceil_div(v->Y, v->TY);
  v->nnz = sp->nnz;
  VGeom g;
  g.d1 = (uint32_t)d[1];
  g.d2 = (uint32_t)d[2];
  g.d3 = (uint32_t)d[3];
  g.rh = v->pos[0];
  g.rl = v->pos[1];
  g.yh = v->pos[2];
  g.yl = v->pos[3];
  g.Rl = (uint32_t)v->Rl;
  g.Yl = (uint32_t)v->Yl;
  g.TY = (uint32_t)v->TY;
  const bool grouped = v->dtype == DT_BF16 || v->dtype == DT_BF16X1;
  g.R = (uint32_t)v->R;
  void *p_slot, *p_cnt, *p_cc;
  const size_t nslot = (size_t)v->ntiles * g.R;
  const uint32_t nch = (uint32_t)std::max<int64_t>(1, ceil_div((int64_t)g.R, (int64_t)VCH));
  FRICC_TRY(ws_get(20, nslot * 4, &p_slot));
  FRICC_TRY(ws_get(21, (size_t)(v->ntiles + 1) * 4, &p_cnt));
  FRICC_TRY(ws_get(23, (size_t)v->ntiles * nch * 4, &p_cc));
  uint32_t* slot = (uint32_t*)p_slot;
  uint32_t* cnt = (uint32_t*)p_cnt;
  uint32_t* cc = (uint32_t*)p_cc;
  if (!v->tile_off) FRICC_HIP(hipMalloc(&v->tile_off, (size_t)(v->ntiles + 1) * 4));
  if (!v->tile_order) FRICC_HIP(hipMalloc(&v->tile_order, (size_t)std::max<int64_t>(v->ntiles, 1) * 4));
  const int64_t nnz = sp->nnz;
  // bytes per gathered row (BF16 views: per plane)
  const size_t rowbytes = (size_t)v->TY * (v->dtype == DT_F64 ? 8 : v->dtype == DT_F32 ? 4 : 2);
  const bool lo_plane = v->dtype == DT_BF16;
  const size_t cap = view_capacity(v->ntiles, g.R, nnz, v->kstep);
  if (v->cap_rows < cap || !v->rowlist) {
    if (v->rowlist) FRICC_HIP(hipFree(v->rowlist));
    if (v->B) FRICC_HIP(hipFree(v->B));
    if (v->Blo) FRICC_HIP(hipFree(v->Blo));
    v->rowlist = nullptr;
    v->B = nullptr;
    v->Blo = nullptr;
    FRICC_HIP(hipMalloc(&v->rowlist, cap * 4));
    FRICC_HIP(hipMalloc(&v->B, cap * rowbytes));
    if (lo_plane) FRICC_HIP(hipMalloc(&v->Blo, cap * rowbytes));
    v->cap_rows = cap;
  }
  k_zero_u32<<<(unsigned)std::min<size_t>(2048, (nslot + VNT - 1) / VNT), VNT, 0, s>>>(slot, (int64_t)nslot);
  if (nnz > 0) k_mark<<<ceil_div(nnz, VNT), VNT, 0, s>>>(sp->idx, nnz, g, slot);
  FRICC_CHECK(v->ntiles < 65536, "view: too many tiles");
  const dim3 gch(nch, (unsigned)v->ntiles);
  k_tile_count<<<gch, VNT, 0, s>>>(slot, g.R, nch, cc);
  k_tile_total<<<(unsigned)ceil_div(v->ntiles, (int64_t)VNT), VNT, 0, s>>>(cc, nch, v->ntiles, v->kstep, cnt);
  FRICC_TRY(scan_u32(cnt, v->tile_off, v->ntiles, v->tile_off + v->ntiles, s));
  k_tile_slots<<<gch, VNT, 0, s>>>(slot, g.R, nch, cc, v->tile_off, v->rowlist);
  k_tile_order<<<(unsigned)std::min<int64_t>(ceil_div(v->ntiles, VNT), 64), VNT, 0, s>>>(v->tile_off, v->ntiles,
                                                                                        v->tile_order);
  const uint32_t* total = v->tile_off + v->ntiles;  // gathered rows, on the device
  if (!grouped) {
    k_zero_rows<<<1024, VNT, 0, s>>>(total, (int64_t)(rowbytes / 16), (uint4*)v->B);
    if (nnz > 0) {
      if (v->dtype == DT_F64)
        k_fill<double><<<ceil_div(nnz, VNT), VNT, 0, s>>>(sp->idx, sp->val, nnz, g, slot, sp->dup_dev,
                                                           sp->dup_host, (double*)v->B);
      else
        k_fill<float><<<ceil_div(nnz, VNT), VNT, 0, s>>>(sp->idx, sp->val, nnz, g, slot, sp->dup_dev,
                                                          sp->dup_host, (float*)v->B);
    }
  } else {
    const int64_t w = v->TY;  // FP32 values per gathered row
    void* p_tmp;
    FRICC_TRY(ws_get(22, cap * (size_t)w * 4, &p_tmp));
    k_zero_rows<<<1024, VNT, 0, s>>>(total, w / 4, (uint4*)p_tmp);
    if (nnz > 0)
      k_fill_grouped<<<ceil_div(nnz, VNT), VNT, 0, s>>>(sp->idx, sp->val, nnz, g, slot, sp->dup_dev,
                                                         sp->dup_host, (float*)p_tmp);
    k_split_rows<<<2048, VNT, 0, s>>>(total, w, (const float*)p_tmp, (__hip_bfloat16*)v->B, (__hip_bfloat16*)v->Blo);
  }
  FRICC_HIP(hipGetLastError());
  return OK;
}

int get_view(Sparse* sp, const int pos[4], int TY, int dtype, hipStream_t s, View** out) {
  auto key = std::make_tuple(pos[0], pos[1], pos[2], pos[3], TY, dtype);
  auto it = sp->views.find(key);
  View* v = nullptr;
  if (it != sp->views.end()) {
    v = it->second.get();
    if (v->gen == sp->gen) {
      *out = v;
      return OK;
    }
  } else {
    std::unique_ptr<View> nv(new View());
    for (int q = 0; q < 4; q++) nv->pos[q] = pos[q];
    nv->TY = TY;
    nv->kstep = (dtype == DT_BF16 || dtype == DT_BF16X1) ? 32 : 16;  // rows per MFMA K step
    nv->dtype = dtype;
    v = nv.get();
    sp->views[key] = std::move(nv);
  }
  FRICC_TRY(build_view(sp, v, s));
  v->gen = sp->gen;
  *out = v;
  return OK;
}

}  // namespace fricc
