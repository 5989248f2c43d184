// Error state, device workspace arena and a generic exclusive u32 scan.
#include "arena.h"
#include "common.h"

#include <mutex>

namespace fricc {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

// Workspace arenas: per device, NARENA independent sets of grow-only slots
// (arena.h).  Work that may run concurrently (solves on different streams)
// selects different arenas; everything recorded or replayed for one solve
// uses that solve's arena, so two solves never share scratch.  The selection
// is per host thread.
struct HipAlloc {
  int alloc(void** p, size_t b) { return (int)hipMalloc(p, b); }
  int free(void* p) { return (int)hipFree(p); }
  int sync() { return (int)hipDeviceSynchronize(); }
  const char* error_string(int e) { return hipGetErrorString((hipError_t)e); }
};

constexpr int MAXDEV = 16;
static ArenaBook<HipAlloc> g_book[MAXDEV];
static std::mutex g_ws_mu;
static thread_local int g_arena = 0;

static int arena_status(int st) {
  return st == ARENA_OK ? OK : st == ARENA_EINVAL ? EINVAL_ : st == ARENA_ENOMEM ? ENOMEM_ : EDEVICE;
}

static int current_device(int* dev) {
  FRICC_HIP(hipGetDevice(dev));
  if (*dev < 0 || *dev >= MAXDEV) {
    set_error("workspace: device index out of range");
    return EINVAL_;
  }
  return OK;
}

int ws_select(int arena) {
  if (!ArenaBook<HipAlloc>::valid_arena(arena)) {
    set_error("workspace arena out of range (0..63)");
    return EINVAL_;
  }
  g_arena = arena;
  return OK;
}

int ws_release(int arena) {
  int dev = 0;
  FRICC_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  HipAlloc al;
  std::string err;
  const int st = g_book[dev].release(al, arena, &err);
  if (st != ARENA_OK) set_error(err);
  return arena_status(st);
}

int ws_pin(int arena, int delta) {
  int dev = 0;
  FRICC_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  std::string err;
  const int st = g_book[dev].pin(arena, delta, &err);
  if (st != ARENA_OK) set_error(err);
  return arena_status(st);
}

// Grow-only slot of the selected arena.  Contents are not preserved on growth.
int ws_get(int slot, size_t bytes, void** out) {
  int dev = 0;
  FRICC_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  HipAlloc al;
  std::string err;
  const int st = g_book[dev].get(al, g_arena, slot, bytes, out, &err);
  if (st != ARENA_OK) set_error(err);
  return arena_status(st);
}

// --- exclusive scan of u32 ----------------------------------------------------
constexpr int SCAN_NT = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_NT * SCAN_ITEMS;

__global__ void k_scan_tile_sums(const uint32_t* __restrict__ in, int64_t n, uint32_t* __restrict__ sums) {
  __shared__ uint32_t sm[SCAN_NT / WAVE + 1];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
  for (int r = 0; r < SCAN_ITEMS; r++) {
    int64_t i = base + (int64_t)r * SCAN_NT + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  block_excl_scan_u32<SCAN_NT>(s, &tot, sm);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Single block: exclusive scan of the tile sums in place; writes the total.
__global__ void k_scan_sums(uint32_t* sums, int64_t nt, uint32_t* total) {
  __shared__ uint32_t sm[1024 / WAVE + 1];
  const int64_t chunk = (nt + 1023) / 1024;
  const int64_t lo = threadIdx.x * chunk, hi = lo + chunk < nt ? lo + chunk : nt;
  uint32_t s = 0;
  for (int64_t i = lo; i < hi; i++) s += sums[i];
  uint32_t tot;
  uint32_t ex = block_excl_scan_u32<1024>(s, &tot, sm);
  for (int64_t i = lo; i < hi; i++) {
    uint32_t t = sums[i];
    sums[i] = ex;
    ex += t;
  }
  if (threadIdx.x == 0 && total) *total = tot;
}

__global__ void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, int64_t n,
                             const uint32_t* __restrict__ sums) {
  __shared__ uint32_t sm[SCAN_NT / WAVE + 1];
  // blocked arrangement per thread: SCAN_ITEMS consecutive elements
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < SCAN_ITEMS; r++) {
    int64_t i = base + r;
    v[r] = i < n ? in[i] : 0u;
    s += v[r];
  }
  uint32_t tot;
  uint32_t ex = block_excl_scan_u32<SCAN_NT>(s, &tot, sm) + sums[blockIdx.x];
#pragma unroll
  for (int r = 0; r < SCAN_ITEMS; r++) {
    int64_t i = base + r;
    if (i < n) out[i] = ex;
    ex += v[r];
  }
}

int64_t scan_u32_sums_words(int64_t n) { return std::max<int64_t>(1, ceil_div(n, SCAN_TILE)); }

int scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev, hipStream_t s, uint32_t* sums) {
  if (n <= 0) return OK;
  const int64_t nt = ceil_div(n, SCAN_TILE);
  if (!sums) {
    void* p = nullptr;
    FRICC_TRY(ws_get(63, (size_t)nt * 4, &p));
    sums = (uint32_t*)p;
  }
  k_scan_tile_sums<<<nt, SCAN_NT, 0, s>>>(in, n, sums);
  k_scan_sums<<<1, 1024, 0, s>>>(sums, nt, total_dev);
  k_scan_tiles<<<nt, SCAN_NT, 0, s>>>(in, out, n, sums);
  FRICC_HIP(hipGetLastError());
  return OK;
}

// Plain streaming copy, 16 bytes per lane per step: the reference point for
// achievable HBM bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE
// counters of the other kernels.
__global__ void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

int stream_copy(const void* src, void* dst, int64_t bytes, hipStream_t s) {
  FRICC_CHECK(bytes >= 0 && bytes % 16 == 0, "stream_copy: bytes must be a multiple of 16");
  FRICC_CHECK(((uintptr_t)src | (uintptr_t)dst) % 16 == 0, "stream_copy: pointers must be 16-byte aligned");
  if (bytes == 0) return OK;
  const int64_t n16 = bytes / 16;
  const int64_t grid = std::min<int64_t>(ceil_div(n16, 256), 256 * 32);
  k_stream_copy<<<grid, 256, 0, s>>>((const uint4*)src, (uint4*)dst, n16);
  FRICC_HIP(hipGetLastError());
  return OK;
}

}  // namespace fricc

extern "C" {
// Workspace arena of the following calls on this device (0..63, default 0).
// No counterpart in the reference.
int fricc_workspace_select(int arena) { return fricc::ws_select(arena); }
// Frees an arena's buffers (waits for the device first).
int fricc_workspace_release(int arena) { return fricc::ws_release(arena); }
// +1: a recorded graph holds the arena's slots; -1: that graph was released.
int fricc_workspace_pin(int arena, int delta) { return fricc::ws_pin(arena, delta); }

// Device-to-device copy of `bytes` (a multiple of 16, 16-byte aligned
// pointers) by a plain streaming kernel.  Bandwidth reference for the
// roofline; no counterpart in the reference.
int fricc_stream_copy(const void* src, void* dst, int64_t bytes, void* stream) {
  return fricc::stream_copy(src, dst, bytes, (hipStream_t)stream);
}
}
