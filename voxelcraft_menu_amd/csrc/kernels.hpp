// kernels.hpp -- kernel declarations shared by the host translation units.
#pragma once

#include "march_core.hpp"

namespace vxd {

// blockIdx -> work block so that consecutive work blocks land on one XCD (one L2):
// the dispatcher hands block b to XCD b % 8; grids are padded to a multiple of 8.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    uint32_t per = nb >> 3;
    return (b & 7u) * per + (b >> 3);
}

__device__ __forceinline__ uint64_t ballot(bool v) { return __ballot(v); }

// Counters are produced by earlier kernels of the same frame; read them with an
// agent-scope atomic load so the value comes from L2, never from a scalar-cache line.
__device__ __forceinline__ uint32_t ld_counter(const uint32_t *c, uint32_t i) {
    return __hip_atomic_load(c + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Control words (the reference's counters[0..5] plus ours) used by the pairs path.
constexpr uint32_t CNT_PAIRS = 0, CNT_DEFERRED = 1, CNT_RECOVER = 2, CNT_SELECT = 3, CNT_SELECT_RAW = 4,
                   CNT_RECOVER_RAW = 5, CNT_SHADOW = 6, CNT_TRUNC = 7;
// Statistics accumulated by every tile's wave go to one of 64 slots, each on its own
// 128-byte line, so tens of thousands of end-of-wave atomics do not serialise on one
// address; readback sums the slots.
constexpr uint32_t STAT_SELECT = 0, STAT_DEFERRED = 1, STAT_RECOVER = 2, STAT_SHADOW = 3;
constexpr uint32_t CNT_STAT_BASE = 32, CNT_STAT_SLOTS = 64, CNT_STAT_STRIDE = 32;
constexpr uint32_t CNT_LEN = CNT_STAT_BASE + CNT_STAT_SLOTS * CNT_STAT_STRIDE;

__device__ __forceinline__ void stat_add(uint32_t *counters, uint32_t slot, uint32_t k, uint32_t v) {
    if (v) atomicAdd(&counters[CNT_STAT_BASE + (slot & (CNT_STAT_SLOTS - 1)) * CNT_STAT_STRIDE + k], v);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n);
__global__ void k_fill_u64(uint64_t *p, uint64_t v, size_t n);
__global__ void k_tile_select(Pools p, FrameK f, const uint32_t *__restrict__ hiz, uint32_t *__restrict__ pairs,
                              uint32_t *__restrict__ counters);
__global__ void k_finalize(FrameK f, uint32_t *counters);
__global__ void k_march_pairs(Pools p, FrameK f, Spec s, const uint32_t *__restrict__ pairs,
                              const uint32_t *__restrict__ counters, unsigned long long *__restrict__ vis);
__global__ void k_build_hiz(FrameK f, const unsigned long long *__restrict__ vis, uint32_t *__restrict__ hiz);
__global__ void k_recover_select(Pools p, FrameK f, const uint32_t *__restrict__ hiz, uint32_t *__restrict__ pairs,
                                 uint32_t *__restrict__ counters);
__global__ void k_finalize_recover(FrameK f, uint32_t *counters);
// Tile-march structure bits (vx_tuning.variant).
//   VAR_SPLIT: classify the whole chunk list into LDS bitmasks before marching; otherwise
//              each 64-chunk block is classified and marched in turn.
//   VAR_FUSED_SHADOW: trace each pixel's sun ray in the tile's wave right after its key.
constexpr int VAR_SPLIT = 1, VAR_FUSED_SHADOW = 4;

template <int kOcc, int kVar, bool kUseHiz, bool kUnderwater>
__global__ void k_march_tiles(Pools p, FrameK f, Spec s, uint32_t *__restrict__ hiz,
                              unsigned long long *__restrict__ vis, uint16_t *__restrict__ shadow,
                              uint32_t *__restrict__ counters, const uint32_t *__restrict__ order,
                              uint32_t *__restrict__ cost);
__global__ void k_order_tiles(const uint32_t *__restrict__ cost, uint32_t *__restrict__ order, uint32_t n);
__global__ void k_iota(uint32_t *__restrict__ p, uint32_t n);
__global__ void k_shadow(Pools p, FrameK f, Spec s, const unsigned long long *__restrict__ vis,
                         uint16_t *__restrict__ out, uint32_t *__restrict__ counters);
__global__ void k_decode(Pools p, FrameK f, const unsigned long long *__restrict__ vis,
                         const uint16_t *__restrict__ shadow, vx_hit *__restrict__ out);

}  // namespace vxd
