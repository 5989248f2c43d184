// Global->LDS DMA helpers shared by the LDS-pipelined kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace fricc {

typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) void gbl_void;

// 16 bytes per lane from global memory straight into LDS: lane l of the wave
// writes LDS byte address lds + 16*l (lds the same for the whole wave).
// Written as one asm block (M0 set, the one wait state M0 needs before an LDS
// DMA, the DMA) so that the compiler does not serialise the pipeline with
// waits of its own; the block has no register outputs, and completion is
// awaited explicitly with wait_vm before the data is read.
__device__ inline void dma16(const void* g, uint32_t lds) {
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g),
               "s"(__builtin_amdgcn_readfirstlane(lds))
               : "m0");
}

__device__ inline uint32_t lds_addr(const void* p) { return (uint32_t)(uintptr_t)(lds_void*)p; }

// s_waitcnt vmcnt(N), other counters untouched (gfx9 encoding)
template <int N>
__device__ inline void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

}  // namespace fricc
