// tiles.hip -- the tile-per-wave primary march (VX_PATH_TILES) and its work ordering.
//
// One 64-lane workgroup owns one 8x8 tile for the whole frame:
//   1. lane j runs the tile_select tests (tile_select.wgsl:11-92) on chunk base+j, the
//      wave ballots the selected / Hi-Z-deferred sets;
//   2. selected chunks are marched in list order (the list is distance sorted), each
//      lane skipping a chunk once its key beats every hit that chunk can produce;
//   3. the tile's Hi-Z word is the wave minimum of the 24-bit depths (build_hiz,
//      march.wgsl:57-76), exactly the reference value because pruning never changes a
//      pixel's maximum key; deferred chunks with depth_key(dmin) >= that word are
//      recovered (recover_select, tile_select.wgsl:94-110) and marched;
//   4. one store per pixel; optionally the pixel's sun-shadow ray (shadow.wgsl:34-48)
//      runs in the same wave on the key it just produced.
// Workgroups are one wave so a long tile never holds idle slots of other waves, and the
// launch order is a list sorted by each tile's measured cost in the previous frame,
// most expensive first (k_order_tiles). Order only schedules: results do not depend on it.
#include "kernels.hpp"

namespace vxd {

// Conservative lower bound of the depth key a chunk at box distance dmin can produce:
// every hit point lies in the chunk's box, so its distance is at least dmin; the
// margin covers f32 rounding of the march's entry and step arithmetic.
__device__ __forceinline__ uint32_t prune_key(float far_, float dmin) {
    float lo = dmin - (0.25f + 1e-3f * dmin);
    return depth_key(far_, fmaxf(lo, 0.0f));
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool kUseHiz, bool kUnderwater, int kVar>
__device__ __forceinline__ void march_tile_body(const Pools &p, const FrameK &f, const Spec &s, uint32_t tile,
                                                uint32_t *__restrict__ hiz, unsigned long long *__restrict__ vis,
                                                uint16_t *__restrict__ shadow, uint32_t *__restrict__ counters,
                                                uint64_t *defmask, uint64_t *selmask) {
    const uint32_t lane = lane_id();
    const uint32_t gx = tile % f.tiles[0], gy = tile / f.tiles[0];
    const uint32_t px = gx * 8u + (lane & 7u), py = gy * 8u + (lane >> 3);
    const bool in_frame = px < f.res[0] && py < f.res[1];
    float rd[3], inv[3];
    ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd[0], rd[1], rd[2]);
#pragma unroll
    for (int i = 0; i < 3; i++) inv[i] = 1.0f / rd[i];
    TileCull tc = tile_cull_setup(f, gx, gy);
    const uint32_t tile_far = kUseHiz ? hiz[tile] : 0u;
    uint64_t best = 0;
    uint32_t n_sel = 0, n_def = 0, n_rec = 0;
    const uint32_t nwords = (f.chunk_count + 63u) >> 6;

    auto march_one = [&](uint32_t cj, uint32_t pk) {
        // pk: prune key of chunk cj; a lane whose key already beats every hit the chunk
        // can produce (pk < best depth) does not march it.
        bool need = in_frame && pk >= (uint32_t)(best >> 40);
        if (!ballot(need)) return;
        ChunkK c = ld_chunk(p, cj);
        if (need && !ray_skips(f, c.fade, px, py)) {
            Hit h = march_chunk<kUnderwater, false, false>(p, f, s, c, f.cam, rd, inv, f.far_);
            if (h.hit) {
                uint64_t k = make_key(f.far_, h, cj);
                best = k > best ? k : best;
            }
        }
    };

    if constexpr ((kVar & VAR_SPLIT) != 0) {
        // Classify the whole chunk list first: the box loads of all words are independent.
        for (uint32_t w = 0; w < nwords; w++) {
            uint32_t ci = (w << 6) + lane;
            uint32_t cls = 0;
            float dmin = 0.0f;
            if (ci < f.chunk_count && ci < p.n_chunks) {
                ChunkK c = ld_chunk(p, ci);
                cls = classify(f, tc, c.bmin, c.bmax, ci, tile_far, kUseHiz, dmin);
            }
            uint64_t sel = ballot(cls == 1u);
            uint64_t def = ballot(cls == 2u);
            n_sel += (uint32_t)__popcll(sel);
            if (lane == 0) selmask[w] = sel;
            if (kUseHiz) {
                n_def += (uint32_t)__popcll(def);
                if (lane == 0) defmask[w] = def;
            }
        }
        wave_lds_sync();
        for (uint32_t w = 0; w < nwords; w++) {
            uint64_t sel = selmask[w];
            if (!sel) continue;
            uint32_t pk = 0;
            if ((sel >> lane) & 1ull) {
                ChunkK c = ld_chunk(p, (w << 6) + lane);
                pk = prune_key(f.far_, box_distance(f.cam, c.bmin, c.bmax));
            }
            while (sel) {
                uint32_t j = (uint32_t)__ffsll((long long)sel) - 1u;
                sel &= sel - 1ull;
                march_one((w << 6) + j, (uint32_t)__shfl((int)pk, (int)j));
            }
        }
    } else {
        for (uint32_t w = 0; w < nwords; w++) {
            uint32_t ci = (w << 6) + lane;
            uint32_t cls = 0;
            float dmin = 0.0f;
            if (ci < f.chunk_count && ci < p.n_chunks) {
                ChunkK c = ld_chunk(p, ci);
                cls = classify(f, tc, c.bmin, c.bmax, ci, tile_far, kUseHiz, dmin);
            }
            uint32_t pk = prune_key(f.far_, dmin);
            uint64_t sel = ballot(cls == 1u);
            uint64_t def = ballot(cls == 2u);
            n_sel += (uint32_t)__popcll(sel);
            if (kUseHiz) {
                n_def += (uint32_t)__popcll(def);
                if (lane == 0) defmask[w] = def;
            }
            while (sel) {
                uint32_t j = (uint32_t)__ffsll((long long)sel) - 1u;
                sel &= sel - 1ull;
                march_one((w << 6) + j, (uint32_t)__shfl((int)pk, (int)j));
            }
        }
    }
    if (kUseHiz) {
        // defmask was written by lane 0 of this wave; LDS operations of a wave complete in order.
        wave_lds_sync();
        uint32_t d = in_frame ? (uint32_t)(best >> 40) : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) d = min(d, (uint32_t)__shfl_xor((int)d, o));
        if (lane == 0) hiz[tile] = d;
        for (uint32_t w = 0; w < nwords; w++) {
            uint64_t def = defmask[w];
            if (!def) continue;
            uint32_t ci = (w << 6) + lane;
            bool rec = false;
            uint32_t pk = 0;
            if ((def >> lane) & 1ull) {
                ChunkK c = ld_chunk(p, ci);
                float dmin = box_distance(f.cam, c.bmin, c.bmax);
                rec = depth_key(f.far_, dmin) >= d;
                pk = prune_key(f.far_, dmin);
            }
            uint64_t r = ballot(rec);
            n_rec += (uint32_t)__popcll(r);
            while (r) {
                uint32_t j = (uint32_t)__ffsll((long long)r) - 1u;
                r &= r - 1ull;
                march_one((w << 6) + j, (uint32_t)__shfl((int)pk, (int)j));
            }
        }
    } else {
        if (lane == 0) hiz[tile] = 0u;
    }
    const size_t pix = (size_t)py * f.res[0] + px;
    if (in_frame) vis[pix] = best;
    uint32_t n_shadow = 0;
    if constexpr ((kVar & VAR_FUSED_SHADOW) != 0) {
        // primary_shadow reads only its own pixel's key: trace it here from the register.
        bool traced = false;
        if (in_frame) shadow[pix] = shadow_pixel(p, f, s, best, px, py, traced);
        n_shadow = (uint32_t)__popcll(ballot(traced));
    }
    if (lane == 0) {
        stat_add(counters, tile, STAT_SELECT, n_sel);
        stat_add(counters, tile, STAT_DEFERRED, n_def);
        stat_add(counters, tile, STAT_RECOVER, n_rec);
        stat_add(counters, tile, STAT_SHADOW, n_shadow);
    }
}

constexpr uint32_t kPrioTiles = 256;

template <int kOcc, int kVar, bool kUseHiz, bool kUnderwater>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kOcc))) void k_march_tiles(
    Pools p, FrameK f, Spec s, uint32_t *__restrict__ hiz, unsigned long long *__restrict__ vis,
    uint16_t *__restrict__ shadow, uint32_t *__restrict__ counters, const uint32_t *__restrict__ order,
    uint32_t *__restrict__ cost) {
    __shared__ uint64_t defmask[kUseHiz ? 128 : 1];
    __shared__ uint64_t selmask[(kVar & VAR_SPLIT) ? 128 : 1];
    const uint32_t ntiles = f.tiles[0] * f.tiles[1];
    if (blockIdx.x >= ntiles) return;
    const uint32_t tile = order[blockIdx.x];
    if (tile >= ntiles) return;
    // The order list starts with last frame's most expensive tiles: they are the
    // frame's critical path, so their waves get instruction-issue priority.
    if (blockIdx.x < kPrioTiles) __builtin_amdgcn_s_setprio(3);
    const uint64_t t0 = wall_clock64();
    march_tile_body<kUseHiz, kUnderwater, kVar>(p, f, s, tile, hiz, vis, shadow, counters, defmask, selmask);
    if (lane_id() == 0) {
        uint64_t dt = wall_clock64() - t0;
        cost[tile] = dt > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dt;
    }
}

#define VX_TILES_INST1(O, V, H, U)                                                                     \
    template __global__ void k_march_tiles<O, V, H, U>(Pools, FrameK, Spec, uint32_t *,                \
                                                       unsigned long long *, uint16_t *, uint32_t *, \
                                                       const uint32_t *, uint32_t *);
#define VX_TILES_INST(O, V)          \
    VX_TILES_INST1(O, V, false, false) \
    VX_TILES_INST1(O, V, true, false)  \
    VX_TILES_INST1(O, V, false, true)  \
    VX_TILES_INST1(O, V, true, true)
VX_TILES_INST(4, 0)
VX_TILES_INST(4, 1)
VX_TILES_INST(4, 4)
VX_TILES_INST(4, 5)
VX_TILES_INST(6, 0)
VX_TILES_INST(6, 4)

// ---- Batched march: many views, one launch -------------------------------------------
// A queue entry names (view, tile). The view's frame, spec and buffer offsets are read
// from its descriptor (wave-uniform, so scalar loads), and the tile runs the same body
// as the single-view kernel, so a pixel's key, shadow word and Hi-Z word are the same
// bits whichever way the views are launched.
template <int kVar>
__device__ __forceinline__ void batch_tile(const BatchArgs &a, uint32_t entry, uint32_t n_views, uint64_t *defmask,
                                           uint64_t *selmask) {
    entry = (uint32_t)__builtin_amdgcn_readfirstlane((int)entry);  // uniform: descriptor reads are scalar
    const uint32_t v = entry >> ENTRY_VIEW_SHIFT, tile = entry & ENTRY_TILE_MASK;
    if (v >= n_views) return;
    // References into the descriptor, not copies: the march indexes frame arrays with
    // run-time indices, and a local copy would be placed in per-lane scratch memory.
    const ViewK *vk = a.views + v;
    const FrameK &f = vk->f;
    const Spec &s = vk->s;
    if (tile >= f.tiles[0] * f.tiles[1]) return;
    Pools p = a.p;
    p.chunks = a.p.chunks + vk->chunk_off;
    p.grid = a.p.grid + vk->grid_off;
    p.n_chunks = vk->n_chunks;
    const size_t px0 = (size_t)vk->px_off_lo | ((size_t)vk->px_off_hi << 32);
    const uint32_t toff = vk->tile_off;
    uint32_t *hiz = a.hiz + toff;
    unsigned long long *vis = a.vis + px0;
    uint16_t *shadow = a.shadow + px0;
    const uint64_t t0 = wall_clock64();
    const bool hz = (f.flags & VX_FLAG_HIZ) != 0u, uw = (f.flags & VX_FLAG_UNDERWATER) != 0u;
    if (uw) {
        if (hz) march_tile_body<true, true, kVar>(p, f, s, tile, hiz, vis, shadow, a.counters, defmask, selmask);
        else march_tile_body<false, true, kVar>(p, f, s, tile, hiz, vis, shadow, a.counters, defmask, selmask);
    } else {
        if (hz) march_tile_body<true, false, kVar>(p, f, s, tile, hiz, vis, shadow, a.counters, defmask, selmask);
        else march_tile_body<false, false, kVar>(p, f, s, tile, hiz, vis, shadow, a.counters, defmask, selmask);
    }
    if (lane_id() == 0) {
        uint64_t dt = wall_clock64() - t0;
        a.cost[toff + tile] = dt > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dt;
    }
}

template <int kOcc, int kVar>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kOcc))) void k_batch_flat(BatchArgs a) {
    __shared__ uint64_t defmask[128];
    __shared__ uint64_t selmask[(kVar & VAR_SPLIT) ? 128 : 1];
    const uint32_t n = a.hdr[BH_TILES];
    if (blockIdx.x >= n) return;
    if (blockIdx.x < kPrioTiles) __builtin_amdgcn_s_setprio(3);
    batch_tile<kVar>(a, a.order[blockIdx.x], a.hdr[BH_VIEWS], defmask, selmask);
}

// Persistent form: the grid is sized to the resident wave slots; every wave takes
// `chunk` consecutive queue entries per atomic on the head word until the queue is
// exhausted (the head only grows, so every wave reaches the exit). The queue is sorted
// most expensive first, so small chunks keep the heavy head of the queue spread over
// many waves.
template <int kOcc, int kVar>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kOcc))) void k_batch_persistent(BatchArgs a,
                                                                                                   uint32_t chunk) {
    __shared__ uint64_t defmask[128];
    __shared__ uint64_t selmask[(kVar & VAR_SPLIT) ? 128 : 1];
    const uint32_t n = a.hdr[BH_TILES];
    const uint32_t nv = a.hdr[BH_VIEWS];
    uint32_t *head = a.counters + CNT_CTL + CTL_HEAD;
    const uint32_t grab = chunk ? chunk : 1u;
    for (;;) {
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(head, grab);
        base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);  // all lanes active: lane 0's value
        if (base >= n) break;
        if (base < kPrioTiles) __builtin_amdgcn_s_setprio(3);
        else __builtin_amdgcn_s_setprio(0);
        const uint32_t end = min(base + grab, n);
        for (uint32_t k = base; k < end; k++) batch_tile<kVar>(a, a.order[k], nv, defmask, selmask);
    }
}

#define VX_BATCH_INST(O, V)                                        \
    template __global__ void k_batch_flat<O, V>(BatchArgs);        \
    template __global__ void k_batch_persistent<O, V>(BatchArgs, uint32_t);
VX_BATCH_INST(4, 4)
VX_BATCH_INST(4, 5)
VX_BATCH_INST(6, 4)

// Launch order for the next frame: tiles grouped into 32 buckets by the log2 of their
// measured cycles, most expensive bucket first (longest-processing-time-first list
// scheduling: the heavy tiles are the frame's critical path, so they start first).
// k_cost_hist counts the buckets over the whole grid; k_cost_scatter gives every block a
// range inside each bucket with one atomic per bucket and places its tiles there. Order
// inside a bucket follows block arrival; it only schedules work, results do not depend
// on it.
__device__ __forceinline__ uint32_t cost_bucket(uint32_t c) { return 31u - (uint32_t)__clz(c | 1u); }

// Entry of global tile i: i itself, or (view, tile-in-view) found in the view table.
__device__ __forceinline__ uint32_t tile_entry(uint32_t i, const uint32_t *toff, uint32_t nv) {
    if (nv == 0u) return i;
    uint32_t lo = 0, hi = nv;  // toff[lo] <= i < toff[hi]
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (toff[mid] <= i) lo = mid;
        else hi = mid;
    }
    return (lo << ENTRY_VIEW_SHIFT) | (i - toff[lo]);
}

__global__ __launch_bounds__(256) void k_cost_hist(const uint32_t *__restrict__ cost, const uint32_t *__restrict__ n_dev,
                                                   uint32_t n, uint32_t *__restrict__ ctl) {
    __shared__ uint32_t hist[COST_BUCKETS];
    if (n_dev) n = *n_dev;
    const uint32_t t = threadIdx.x;
    if (t < COST_BUCKETS) hist[t] = 0;
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + t; i < n; i += stride) atomicAdd(&hist[cost_bucket(cost[i])], 1u);
    __syncthreads();
    if (t < COST_BUCKETS && hist[t]) atomicAdd(&ctl[CTL_HIST + t], hist[t]);
}

__global__ __launch_bounds__(256) void k_cost_scatter(const uint32_t *__restrict__ cost,
                                                      const uint32_t *__restrict__ n_dev, uint32_t n,
                                                      uint32_t *__restrict__ ctl, const uint32_t *__restrict__ tile_off,
                                                      const uint32_t *__restrict__ n_views_dev,
                                                      uint32_t *__restrict__ order) {
    __shared__ uint32_t cnt[COST_BUCKETS];
    __shared__ uint32_t base[COST_BUCKETS];
    __shared__ uint32_t toff[BATCH_MAX_VIEWS + 1];
    if (n_dev) n = *n_dev;
    const uint32_t nv = n_views_dev ? min(*n_views_dev, BATCH_MAX_VIEWS) : 0u;
    const uint32_t t = threadIdx.x;
    if (t < COST_BUCKETS) cnt[t] = 0;
    if (nv)
        for (uint32_t v = t; v <= nv; v += blockDim.x) toff[v] = tile_off[v];
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + t; i < n; i += stride) atomicAdd(&cnt[cost_bucket(cost[i])], 1u);
    __syncthreads();
    if (t < COST_BUCKETS) {
        uint32_t start = 0;
        for (uint32_t b = t + 1u; b < COST_BUCKETS; b++) start += ctl[CTL_HIST + b];
        const uint32_t c = cnt[t];
        base[t] = c ? start + atomicAdd(&ctl[CTL_CURSOR + t], c) : 0u;
        cnt[t] = 0;
    }
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + t; i < n; i += stride) {
        const uint32_t b = cost_bucket(cost[i]);
        order[base[b] + atomicAdd(&cnt[b], 1u)] = tile_entry(i, toff, nv);
    }
}

__global__ __launch_bounds__(256) void k_iota_views(uint32_t *__restrict__ order, const uint32_t *__restrict__ n_dev,
                                                    uint32_t n, const uint32_t *__restrict__ tile_off,
                                                    const uint32_t *__restrict__ n_views_dev) {
    __shared__ uint32_t toff[BATCH_MAX_VIEWS + 1];
    if (n_dev) n = *n_dev;
    const uint32_t nv = n_views_dev ? min(*n_views_dev, BATCH_MAX_VIEWS) : 0u;
    if (nv)
        for (uint32_t v = threadIdx.x; v <= nv; v += blockDim.x) toff[v] = tile_off[v];
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) order[i] = tile_entry(i, toff, nv);
}

}  // namespace vxd
