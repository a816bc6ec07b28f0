// kernels.hpp -- kernel declarations shared by the host translation units.
#pragma once

#include "march_core.hpp"

namespace vxd {

constexpr uint32_t CNT_PAIRS = 0, CNT_DEFERRED = 1, CNT_RECOVER = 2, CNT_SELECT = 3, CNT_SELECT_RAW = 4,
                   CNT_RECOVER_RAW = 5, CNT_SHADOW = 6, CNT_TRUNC = 7, CNT_LEN = 16;

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
__global__ void k_march_tiles(Pools p, FrameK f, Spec s, uint32_t *__restrict__ hiz,
                              unsigned long long *__restrict__ vis, uint32_t *__restrict__ counters);
__global__ void k_shadow(Pools p, FrameK f, Spec s, const unsigned long long *__restrict__ vis,
                         uint16_t *__restrict__ out, uint32_t *__restrict__ counters);
__global__ void k_decode(Pools p, FrameK f, const unsigned long long *__restrict__ vis,
                         const uint16_t *__restrict__ shadow, vx_hit *__restrict__ out);

}  // namespace vxd
