// secondary.hip -- the secondary bounce of a batch: one leg per water/glass hit and leg
// kind, gathered into compact queues and traced one leg per lane.
//
// Legs follow resolve.wgsl's shade_water (refract to 96, reflect to 768 along water_ctx's
// mirror_t, the wave normal on waved top faces) and shade_glass (transmit, and reflect
// under SPEC_GLASS_REFLECT, to max(192, far)), each traced with trace_world
// (common.wgsl:1205-1253). Only a minority of pixels hit water or glass, so the tile
// waves of the gather pass append the legs they own to two queues (water legs skip water
// voxels, glass legs do not: one trace kernel specialisation each) and the trace pass
// runs full waves of legs. A glass transmit leg that ends on water is shaded by
// shade_water there (resolve.wgsl:1281-1289): the glass trace appends those two water
// legs to a third queue, traced after it from the transmit leg's record.
#include "kernels.hpp"

namespace vxd {

// Output record of a leg (vx_sec_hit): t, voxel, info = ci | normal << 13 | sun blocked
// << 28 | sun tested << 29 | traced << 30 | hit << 31 (ci = 0x1FFF on a miss), block id
// at the hit. The sun test at the hit (leg_sun) runs here, in the lane that traced it.
__device__ __forceinline__ uint4 leg_record(const Pools &p, const FrameK &f, const Spec &s, const SecLeg &l,
                                            const WorldHit &w) {
    uint32_t material = 0u, sun = 0u;
    if (w.hit) {
        const ChunkK c = ld_chunk(p, w.ci);
        material = block_at(p, c, w.voxel & 63u, (w.voxel >> 6) & 63u, (w.voxel >> 12) & 63u);
        sun = leg_sun(p, f, s, l, w, c, material);
    }
    const uint32_t info = (w.hit ? (w.ci & 0x1FFFu) : 0x1FFFu) | ((w.hit ? w.normal : 0u) << 13) | (1u << 30) |
                          (w.hit ? (1u << 31) : 0u) | sun;
    return make_uint4(__float_as_uint(w.t), w.hit ? w.voxel : 0u, info, material);
}

// Appends this lane's entry to queue q (its counter in line `slot`) when `want` (one
// atomic per wave and queue).
__device__ __forceinline__ void queue_push(const SecArgs &s, uint32_t *counters, uint32_t q, uint32_t slot, bool want,
                                           unsigned long long entry) {
    const uint64_t m = ballot(want);
    if (!m) return;
    const uint32_t lane = lane_id();
    const uint32_t first = (uint32_t)__ffsll((long long)m) - 1u;
    uint32_t base = 0;
    if (lane == first) base = atomicAdd(&counters[CNT_SEC + slot * SEC_LINE], (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, (int)first);
    if (want) {
        const uint32_t idx = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (idx < s.cap) s.queue[(size_t)q * s.cap + idx] = entry;
        else counters[CNT_SEC + SEC_OVERFLOW * SEC_LINE] = 1u;
    }
}

// Tiles are gathered in view order (block i is global tile i, its view found by binary
// search of the tile offsets), so each queue holds the legs view after view and the trace
// waves rarely span more than one view.
__global__ __launch_bounds__(64) void k_sec_gather(BatchArgs a, SecArgs s) {
    const uint32_t n = a.hdr[BH_TILES], nv = a.hdr[BH_VIEWS];
    if (blockIdx.x >= n || nv == 0u) return;
    uint32_t lo = 0, hi = nv;  // tile_off[lo] <= blockIdx.x < tile_off[hi]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s.tile_off[mid] <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const uint32_t v = lo, tile = blockIdx.x - s.tile_off[lo];
    const ViewK *vk = a.views + v;
    const FrameK &f = *(const FrameK *)as_const(&vk->f);
    const Spec sp = ld_const(&vk->s);
    if (tile >= f.tiles[0] * f.tiles[1]) return;
    const_as<ViewK> *vc = as_const(vk);
    const Pools &p = *(const Pools *)as_const(&vk->p);
    const size_t px0 = (size_t)vc->px_off_lo | ((size_t)vc->px_off_hi << 32);
    const uint32_t lane = lane_id();
    const uint32_t px = (tile % f.tiles[0]) * 8u + (lane & 7u), py = (tile / f.tiles[0]) * 8u + (lane >> 3);
    const bool in_frame = px < f.res[0] && py < f.res[1];
    const uint32_t pix = in_frame ? py * f.res[0] + px : 0u;
    SecLeg legs[2];
    bool want[2] = {false, false};
    if (in_frame) secondary_legs<false>(p, f, sp, a.vis[px0 + pix], px, py, legs, want);
    for (uint32_t leg = 0; leg < 2; leg++) {
        const bool w = want[leg] && legs[leg].water, g = want[leg] && !legs[leg].water;
        queue_push(s, a.counters, SEC_WATER, SEC_WATER, w, sec_entry(v, pix, leg));
        queue_push(s, a.counters, SEC_GLASS, SEC_GLASS, g, sec_entry(v, pix, leg));
    }
}

// Entry fields: view, pixel, record.
__device__ __forceinline__ uint32_t entry_view(unsigned long long e) { return (uint32_t)(e >> 32); }
__device__ __forceinline__ uint32_t entry_pix(unsigned long long e) { return (uint32_t)(e >> 2) & 0x3FFFFFFFu; }
__device__ __forceinline__ uint32_t entry_rec(unsigned long long e) { return (uint32_t)e & 3u; }

// The previous frame's leg records back to zero: every record written is a queue entry's,
// so clearing the entries of the three queues leaves the whole record array zero. Runs
// before the frame resets the queue counters (a new view set or bounce setting starts
// with zero records and counters).
__global__ __launch_bounds__(256) void k_sec_clear(BatchArgs a, SecArgs s) {
    const uint32_t nv = a.hdr[BH_VIEWS];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t q = 0; q < 3u; q++) {
        const uint32_t slot = q == SEC_Q_SUB ? SEC_SUB : q;
        const uint32_t n = min(ld_counter(a.counters, CNT_SEC + slot * SEC_LINE), s.cap);
        for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
            const unsigned long long e = s.queue[(size_t)q * s.cap + k];
            const uint32_t v = entry_view(e);
            if (v >= nv) continue;
            const ViewK *vk = a.views + v;
            const size_t px0 = (size_t)vk->px_off_lo | ((size_t)vk->px_off_hi << 32);
            s.out[SEC_RECORDS * (px0 + entry_pix(e)) + entry_rec(e)] = make_uint4(0u, 0u, 0u, 0u);
        }
    }
}

// Legs of one queue, one per lane. The lanes of a wave are served view by view (the view
// of the first unserved lane, made uniform), so frame and chunk data stay scalar. The sun
// rays run at the hits are counted per wave and added once at the end. The glass trace
// appends the water legs of transmit legs that end on water to the sub-leg queue.
template <bool kWater>
__global__ __launch_bounds__(64) void k_sec_trace(BatchArgs a, SecArgs s) {
    const uint32_t q = kWater ? SEC_WATER : SEC_GLASS;
    const uint32_t n = min(ld_counter(a.counters, CNT_SEC + q * SEC_LINE), s.cap);
    const uint32_t nv = a.hdr[BH_VIEWS];
    const uint32_t lane = lane_id();
    uint32_t sun_rays = 0;  // wave-uniform
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t k = base + lane;
        const bool mine = k < n;
        const unsigned long long e = mine ? s.queue[(size_t)q * s.cap + k] : 0ull;
        const uint32_t my_view = mine ? entry_view(e) : 0xFFFFFFFFu;
        bool todo = mine && my_view < nv;
        while (ballot(todo) != 0ull) {
            const uint32_t v = (uint32_t)__builtin_amdgcn_readfirstlane(
                (int)__shfl((int)my_view, __ffsll((long long)ballot(todo)) - 1));
            const bool here = todo && my_view == v;
            const ViewK *vk = a.views + v;
            const FrameK &f = *(const FrameK *)as_const(&vk->f);
            const Spec sp = ld_const(&vk->s);
            const_as<ViewK> *vc = as_const(vk);
            const Pools &p = *(const Pools *)as_const(&vk->p);
            const size_t px0 = (size_t)vc->px_off_lo | ((size_t)vc->px_off_hi << 32);
            uint32_t info = 0, pix = 0;
            bool sub[2] = {false, false};
            if (here) {
                pix = entry_pix(e);
                const uint32_t leg = entry_rec(e);
                const uint32_t px = pix % f.res[0], py = pix / f.res[0];
                const uint64_t key = a.vis[px0 + pix];
                SecLeg legs[2];
                bool want[2];
                secondary_legs<true>(p, f, sp, key, px, py, legs, want);
                const SecLeg &l = legs[leg];
                WorldHit w{false, 0.0f, 0u, 0u, NO_CHUNK};
                if (want[leg]) w = trace_world<kWater>(p, f, sp, l.ro, l.rd, l.max_dist);
                const uint4 rec = leg_record(p, f, sp, l, w);
                s.out[SEC_RECORDS * (px0 + pix) + leg] = rec;
                info = rec.z;
                if (!kWater && leg == 0u && w.hit && rec.w == WATER_ID) {
                    SecLeg wl[2];
                    glass_water_legs<false>(p, f, sp, key, px, py, w.t, w.normal, wl, sub);
                }
            }
            sun_rays += (uint32_t)__popcll(ballot((info & SEC_SUN_TESTED) != 0u));
            if constexpr (!kWater) {
                queue_push(s, a.counters, SEC_Q_SUB, SEC_SUB, sub[0], sec_entry(v, pix, 2u));
                queue_push(s, a.counters, SEC_Q_SUB, SEC_SUB, sub[1], sec_entry(v, pix, 3u));
            }
            todo = todo && !here;
        }
    }
    if (lane == 0u && sun_rays) atomicAdd(&a.counters[CNT_SEC + SEC_SUN * SEC_LINE], sun_rays);
}

// The water legs where a glass transmit leg ends on water (records 2 and 3), one per lane,
// from the transmit leg's distance and face in the pixel's record 0 (written by the glass
// trace, which runs before this kernel on the same stream).
__global__ __launch_bounds__(64) void k_sec_trace_sub(BatchArgs a, SecArgs s) {
    const uint32_t n = min(ld_counter(a.counters, CNT_SEC + SEC_SUB * SEC_LINE), s.cap);
    const uint32_t nv = a.hdr[BH_VIEWS];
    const uint32_t lane = lane_id();
    uint32_t sun_rays = 0;  // wave-uniform
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t k = base + lane;
        const bool mine = k < n;
        const unsigned long long e = mine ? s.queue[(size_t)SEC_Q_SUB * s.cap + k] : 0ull;
        const uint32_t my_view = mine ? entry_view(e) : 0xFFFFFFFFu;
        bool todo = mine && my_view < nv;
        while (ballot(todo) != 0ull) {
            const uint32_t v = (uint32_t)__builtin_amdgcn_readfirstlane(
                (int)__shfl((int)my_view, __ffsll((long long)ballot(todo)) - 1));
            const bool here = todo && my_view == v;
            const ViewK *vk = a.views + v;
            const FrameK &f = *(const FrameK *)as_const(&vk->f);
            const Spec sp = ld_const(&vk->s);
            const_as<ViewK> *vc = as_const(vk);
            const Pools &p = *(const Pools *)as_const(&vk->p);
            const size_t px0 = (size_t)vc->px_off_lo | ((size_t)vc->px_off_hi << 32);
            uint32_t info = 0;
            if (here) {
                const uint32_t pix = entry_pix(e), rec = entry_rec(e);
                const uint32_t px = pix % f.res[0], py = pix / f.res[0];
                const uint4 r0 = s.out[SEC_RECORDS * (px0 + pix)];
                SecLeg legs[2];
                bool want[2];
                glass_water_legs<true>(p, f, sp, a.vis[px0 + pix], px, py, __uint_as_float(r0.x), (r0.z >> 13) & 7u,
                                       legs, want);
                const uint32_t leg = rec - 2u;
                const SecLeg &l = legs[leg];
                WorldHit w{false, 0.0f, 0u, 0u, NO_CHUNK};
                if (want[leg]) w = trace_world<true>(p, f, sp, l.ro, l.rd, l.max_dist);
                const uint4 out = leg_record(p, f, sp, l, w);
                s.out[SEC_RECORDS * (px0 + pix) + rec] = out;
                info = out.z;
            }
            sun_rays += (uint32_t)__popcll(ballot((info & SEC_SUN_TESTED) != 0u));
            todo = todo && !here;
        }
    }
    if (lane == 0u && sun_rays) atomicAdd(&a.counters[CNT_SEC + SEC_SUN * SEC_LINE], sun_rays);
}

template __global__ void k_sec_trace<true>(BatchArgs, SecArgs);
template __global__ void k_sec_trace<false>(BatchArgs, SecArgs);

}  // namespace vxd
