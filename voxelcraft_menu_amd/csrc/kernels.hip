// kernels.hip -- gfx950 kernels for the primary march, Hi-Z tile culling and sun shadow.
//
// Two formulations of the primary pass produce the same visibility buffer:
//   * pairs: the reference decomposition (tile_select.wgsl, march.wgsl): a tile culling
//     kernel emits (tile, chunk) pairs, one 64-lane wavefront marches one pair (8x8 tile
//     = one wave64) and merges with a 64-bit atomicMax, then build_hiz / recover_select
//     and a second march.
//   * tiles: one wavefront owns one 8x8 tile for the whole frame. Lane j classifies chunk
//     base+j (the tile_select tests, in registers), the selected chunks are marched in
//     list order (nearest first) with a per-lane exit once a chunk cannot beat the
//     lane's current key, the tile's Hi-Z is the wave minimum after phase one, and the
//     deferred chunks (kept as a bitmask in LDS) are recovered and marched in the same
//     wave. The visibility word is written once, without atomics.
#include "kernels.hpp"

namespace vxd {

// Counter slots (the reference's counters[0..5], src/render/mod.rs:1948-1955, plus ours).
// 0: pairs to march   1: deferred (raw)   2: recover count   3: select (min cap)
// 4: select (raw)     5: recover (raw)    6: shadow rays      7: truncated flag

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u64(uint64_t *p, uint64_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------- pairs path
// tile_select (tile_select.wgsl:11-92): one thread per tile; a wave's lanes walk the
// chunk list in lockstep, so emits are aggregated with one atomic per wave and chunk.
__global__ __launch_bounds__(256) void k_tile_select(Pools p, FrameK f, const uint32_t *__restrict__ hiz,
                                                     uint32_t *__restrict__ pairs, uint32_t *__restrict__ counters) {
    uint32_t gx = blockIdx.x * 16u + (threadIdx.x & 15u);
    uint32_t gy = blockIdx.y * 16u + (threadIdx.x >> 4);
    bool live = gx < f.tiles[0] && gy < f.tiles[1];
    uint32_t tile_id = gy * f.tiles[0] + gx;
    TileCull tc = tile_cull_setup(f, live ? gx : 0u, live ? gy : 0u);
    bool use_hiz = (f.flags & VX_FLAG_HIZ) != 0u;
    uint32_t tile_far = live ? hiz[tile_id] : 0u;
    uint32_t word = tile_id << 13;
    for (uint32_t ci = 0; ci < f.chunk_count; ci++) {
        ChunkK c = ld_chunk(p, ci);
        float dmin;
        uint32_t cls = live ? classify(f, tc, c.bmin, c.bmax, ci, tile_far, use_hiz, dmin) : 0u;
        uint64_t sel = ballot(cls == 1u), def = ballot(cls == 2u);
        uint32_t below = (uint32_t)__popcll(sel & ((1ull << lane_id()) - 1ull));
        uint32_t dbelow = (uint32_t)__popcll(def & ((1ull << lane_id()) - 1ull));
        uint32_t base = 0, dbase = 0;
        if (sel) {
            int leader = __ffsll((long long)sel) - 1;
            if ((int)lane_id() == leader) base = atomicAdd(&counters[0], (uint32_t)__popcll(sel));
            base = (uint32_t)__shfl((int)base, leader);
        }
        if (def) {
            int leader = __ffsll((long long)def) - 1;
            if ((int)lane_id() == leader) dbase = atomicAdd(&counters[1], (uint32_t)__popcll(def));
            dbase = (uint32_t)__shfl((int)dbase, leader);
        }
        if (cls == 1u) {
            uint32_t i = base + below;
            if (i < f.max_pairs) pairs[i] = word | ci;
        } else if (cls == 2u) {
            uint32_t j = dbase + dbelow;
            if (j < f.max_pairs) pairs[f.max_pairs + j] = word | ci;
        }
    }
}

// finalize (tile_select.wgsl:112-122)
__global__ void k_finalize(FrameK f, uint32_t *counters) {
    uint32_t raw = ld_counter(counters, 0);
    counters[4] = raw;
    counters[3] = raw < f.max_pairs ? raw : f.max_pairs;
    if (raw > f.max_pairs || ld_counter(counters, 1) > f.max_pairs) counters[7] = 1u;
}

// march (march.wgsl:1-55): one wave64 per pair, persistent over the pair list.
__global__ __launch_bounds__(256) void k_march_pairs(Pools p, FrameK f, Spec s, const uint32_t *__restrict__ pairs,
                                                     const uint32_t *__restrict__ counters,
                                                     unsigned long long *__restrict__ vis, uint32_t *__restrict__ heat) {
    uint32_t n = ld_counter(counters, 0);
    if (n > f.max_pairs) n = f.max_pairs;
    uint32_t waves = gridDim.x * (blockDim.x >> 6);
    uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    uint32_t lx = threadIdx.x & 7u, ly = (threadIdx.x >> 3) & 7u;
    const bool underwater = (f.flags & VX_FLAG_UNDERWATER) != 0u;
    for (uint32_t q = w; q < n; q += waves) {
        uint32_t pair = pairs[q];
        uint32_t tile = pair >> 13, ci = pair & 0x1FFFu;
        uint32_t px = (tile % f.tiles[0]) * 8u + lx, py = (tile / f.tiles[0]) * 8u + ly;
        if (px >= f.res[0] || py >= f.res[1] || ci >= p.n_chunks) continue;
        ChunkK c = ld_chunk(p, ci);
        if (ray_skips(f, c.fade, px, py)) continue;
        float rd[3], inv[3];
        ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd[0], rd[1], rd[2]);
        for (int i = 0; i < 3; i++) inv[i] = 1.0f / rd[i];
        Hit h = underwater ? march_chunk<true, false, false>(p, f, s, c, ci, f.cam, rd, inv, f.far_)
                           : march_chunk<false, false, false>(p, f, s, c, ci, f.cam, rd, inv, f.far_);
        if (underwater && (s.bits & SPEC_SNELL_BEND) && !h.hit && rd[1] > 0.0f) snell_bend(p, f, s, c, ci, rd, h);
        if (h.hit) atomicMax(&vis[(size_t)py * f.res[0] + px], (unsigned long long)make_key(f.far_, h, ci));
        // FLAG_HEATMAP frames: every marched pair adds its iterations (march.wgsl:51-53).
        if (heat) atomicAdd(&heat[(size_t)py * f.res[0] + px], h.iters);
    }
}

// build_hiz (march.wgsl:57-76): one wave per tile, 0 for pixels off the frame.
__global__ __launch_bounds__(256) void k_build_hiz(FrameK f, const unsigned long long *__restrict__ vis,
                                                   uint32_t *__restrict__ hiz) {
    uint32_t tile = blockIdx.x * 4u + (threadIdx.x >> 6);
    uint32_t ntiles = f.tiles[0] * f.tiles[1];
    if (tile >= ntiles) return;
    uint32_t lane = threadIdx.x & 63u;
    uint32_t px = (tile % f.tiles[0]) * 8u + (lane & 7u), py = (tile / f.tiles[0]) * 8u + (lane >> 3);
    uint32_t d = 0u;
    if (px < f.res[0] && py < f.res[1]) d = (uint32_t)(vis[(size_t)py * f.res[0] + px] >> 40);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) d = min(d, (uint32_t)__shfl_xor((int)d, o));
    if (lane == 0) hiz[tile] = d;
}

// recover_select (tile_select.wgsl:94-110)
__global__ __launch_bounds__(256) void k_recover_select(Pools p, FrameK f, const uint32_t *__restrict__ hiz,
                                                        uint32_t *__restrict__ pairs, uint32_t *__restrict__ counters) {
    uint32_t n = ld_counter(counters, 1);
    if (n > f.max_pairs) n = f.max_pairs;
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    bool rec = false;
    uint32_t pair = 0;
    if (g < n) {
        pair = pairs[f.max_pairs + g];
        uint32_t ci = pair & 0x1FFFu;
        ChunkK c = ld_chunk(p, ci < p.n_chunks ? ci : 0u);
        float dmin = box_distance(f.cam, c.bmin, c.bmax);
        uint32_t tile = pair >> 13;
        uint32_t hz = tile < f.tiles[0] * f.tiles[1] ? hiz[tile] : 0u;
        rec = ci < p.n_chunks && depth_key(f.far_, dmin) >= hz;
    }
    uint64_t m = ballot(rec);
    if (!m) return;
    uint32_t base = 0;
    int leader = __ffsll((long long)m) - 1;
    if ((int)lane_id() == leader) base = atomicAdd(&counters[2], (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader);
    if (rec) {
        uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane_id()) - 1ull));
        if (k < f.max_pairs) pairs[k] = pair;
    }
}

// finalize_recover (tile_select.wgsl:124-133)
__global__ void k_finalize_recover(FrameK f, uint32_t *counters) {
    uint32_t raw = ld_counter(counters, 2);
    counters[5] = raw;
    counters[0] = raw < f.max_pairs ? raw : f.max_pairs;
    if (raw > f.max_pairs) counters[7] = 1u;
}

// ---------------------------------------------------------------- shadow
// primary_shadow (shadow.wgsl:34-48): a wave covers one 8x8 tile.
__global__ __launch_bounds__(256) void k_shadow(Pools p, FrameK f, Spec s, const unsigned long long *__restrict__ vis,
                                                uint16_t *__restrict__ out, uint32_t *__restrict__ counters) {
    uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    uint32_t tile = b * 4u + (threadIdx.x >> 6);
    if (tile >= f.tiles[0] * f.tiles[1]) return;
    uint32_t lane = lane_id();
    uint32_t px = (tile % f.tiles[0]) * 8u + (lane & 7u), py = (tile / f.tiles[0]) * 8u + (lane >> 3);
    bool traced = false;
    if (px < f.res[0] && py < f.res[1]) {
        size_t i = (size_t)py * f.res[0] + px;
        float rd[3];
        ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd[0], rd[1], rd[2]);
        out[i] = shadow_pixel(p, f, s, vis[i], rd, traced);
    }
    uint64_t t = ballot(traced);
    if (lane == 0) stat_add(counters, tile, STAT_SHADOW, (uint32_t)__popcll(t));
}

// ---------------------------------------------------------------- decode
// Visibility key -> (world voxel, face, material, t, shadow bit).
__global__ __launch_bounds__(256) void k_decode(Pools p, FrameK f, const unsigned long long *__restrict__ vis,
                                                const uint16_t *__restrict__ shadow, vx_hit *__restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = f.res[0] * f.res[1];
    if (i >= n) return;
    uint32_t px = i % f.res[0], py = i / f.res[0];
    uint64_t key = vis[i];
    vx_hit o;
    o.voxel[0] = o.voxel[1] = o.voxel[2] = 0;
    o.face = 0;
    o.material = 0;
    o.ci = 0;
    o.t = -1.0f;
    o.shadow = shadow ? (shadow[i] != 0) : 0u;
    uint32_t ci = (uint32_t)(key >> 24) & 0x1FFFu;
    if (key != 0ull && ci < p.n_chunks) {
        uint32_t low = (uint32_t)(key & 0xFFFFFFull);
        uint32_t vx = (low >> 16) & 63u, vy = (low >> 8) & 63u, vz = low & 63u;
        uint32_t mx = (low >> 22) & 3u, my = (low >> 14) & 3u, mz = (low >> 6) & 3u;
        uint32_t nid = (uint32_t)(key >> 37) & 7u;
        ChunkK c = ld_chunk(p, ci);
        float rd[3];
        ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd[0], rd[1], rd[2]);
        o.voxel[0] = (int32_t)c.origin[0] + (int32_t)vx * (int32_t)c.vs;
        o.voxel[1] = (int32_t)c.origin[1] + (int32_t)vy * (int32_t)c.vs;
        o.voxel[2] = (int32_t)c.origin[2] + (int32_t)vz * (int32_t)c.vs;
        o.face = nid;
        o.material = block_at(p, c, vx, vy, vz);
        o.ci = ci;
        o.t = hit_t(f, c, vx, vy, vz, mx, my, mz, nid, rd);
    }
    out[i] = o;
}

}  // namespace vxd
