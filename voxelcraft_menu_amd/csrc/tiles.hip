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
                                                uint64_t *defmask, uint64_t *selmask,
                                                const uint32_t *bin_n = nullptr, const uint16_t *bin_l = nullptr) {
    const uint32_t lane = lane_id();
    const uint32_t gx = tile % f.tiles[0], gy = tile / f.tiles[0];
    const uint32_t px = gx * 8u + (lane & 7u), py = gy * 8u + (lane >> 3);
    const bool in_frame = px < f.res[0] && py < f.res[1];
    float rd[3], inv[3];
    ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd[0], rd[1], rd[2]);
#pragma unroll
    for (int i = 0; i < 3; i++) inv[i] = 1.0f / rd[i];
    TileCull tc = tile_cull_setup_wave(f, gx, gy, lane);
    const uint32_t tile_far = kUseHiz ? hiz[tile] : 0u;
    uint64_t best = 0;
    uint32_t n_sel = 0, n_def = 0, n_rec = 0;
    // The items classified: the view's visible list, or the tile's bin list (the same
    // chunks in the same order, less some that every test of this tile rejects).
    const uint32_t n_items = bin_l ? *bin_n : min(f.chunk_count, p.n_chunks);
    auto chunk_at = [&](uint32_t pos) -> uint32_t { return bin_l ? (uint32_t)bin_l[pos] : pos; };
    const uint32_t nwords = (n_items + 63u) >> 6;

    auto march_one = [&](uint32_t cj, uint32_t pk) {
        // pk: prune key of chunk cj; a lane whose key already beats every hit the chunk
        // can produce (pk < best depth) does not march it.
        bool need = in_frame && pk >= (uint32_t)(best >> 40);
        if (!ballot(need)) return;
        ChunkK c = ld_chunk(p, cj);
        if (need && !ray_skips(f, c.fade, px, py)) {
            Hit h = march_chunk<kUnderwater, false, false>(p, f, s, c, cj, f.cam, rd, inv, f.far_);
            if constexpr (kUnderwater) {
                if ((s.bits & SPEC_SNELL_BEND) && !h.hit && rd[1] > 0.0f) snell_bend(p, f, s, c, cj, rd, h);
            }
            if (h.hit) {
                uint64_t k = make_key(f.far_, h, cj);
                best = k > best ? k : best;
            }
        }
    };

    if constexpr ((kVar & VAR_SPLIT) != 0) {
        // Classify the whole chunk list first: the box loads of all words are independent.
        for (uint32_t w = 0; w < nwords; w++) {
            const uint32_t pos = (w << 6) + lane;
            uint32_t cls = 0;
            float dmin = 0.0f;
            if (pos < n_items) {
                const uint32_t ci = chunk_at(pos);
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
                ChunkK c = ld_chunk(p, chunk_at((w << 6) + lane));
                pk = prune_key(f.far_, box_distance(f.cam, c.bmin, c.bmax));
            }
            while (sel) {
                uint32_t j = (uint32_t)__ffsll((long long)sel) - 1u;
                sel &= sel - 1ull;
                march_one(chunk_at((w << 6) + j), (uint32_t)__shfl((int)pk, (int)j));
            }
        }
    } else {
        for (uint32_t w = 0; w < nwords; w++) {
            const uint32_t pos = (w << 6) + lane;
            uint32_t cls = 0;
            float dmin = 0.0f;
            if (pos < n_items) {
                const uint32_t ci = chunk_at(pos);
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
                march_one(chunk_at((w << 6) + j), (uint32_t)__shfl((int)pk, (int)j));
            }
        }
    }
    if (kUseHiz) {
        // defmask was written by lane 0 of this wave; LDS operations of a wave complete in order.
        wave_lds_sync();
        uint32_t d = in_frame ? (uint32_t)(best >> 40) : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) d = min(d, (uint32_t)__shfl_xor((int)d, o));
        d = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);  // every lane holds the minimum
        if (lane == 0) hiz[tile] = d;
        for (uint32_t w = 0; w < nwords; w++) {
            uint64_t def = defmask[w];
            if (!def) continue;
            bool rec = false;
            uint32_t pk = 0;
            if ((def >> lane) & 1ull) {
                ChunkK c = ld_chunk(p, chunk_at((w << 6) + lane));
                float dmin = box_distance(f.cam, c.bmin, c.bmax);
                rec = depth_key(f.far_, dmin) >= d;
                pk = prune_key(f.far_, dmin);
            }
            uint64_t r = ballot(rec);
            n_rec += (uint32_t)__popcll(r);
            while (r) {
                uint32_t j = (uint32_t)__ffsll((long long)r) - 1u;
                r &= r - 1ull;
                march_one(chunk_at((w << 6) + j), (uint32_t)__shfl((int)pk, (int)j));
            }
        }
    } else {
        if (lane == 0) hiz[tile] = 0u;
    }
    // The lane's pixel again, from the lane id and the tile (both free to recompute), so
    // the coordinates need not be held through the march. Every lane is active here.
    uint32_t lane2;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane2));
    const uint32_t px2 = gx * 8u + (lane2 & 7u), py2 = gy * 8u + (lane2 >> 3);
    const bool in_frame2 = px2 < f.res[0] && py2 < f.res[1];
    const size_t pix = (size_t)py2 * f.res[0] + px2;
    if (in_frame2) vis[pix] = best;
    uint32_t n_shadow = 0;
    if constexpr ((kVar & VAR_FUSED_SHADOW) != 0) {
        // primary_shadow reads only its own pixel's key: trace it here from the register.
        bool traced = false;
        if (in_frame2) shadow[pix] = shadow_pixel(p, f, s, best, rd, traced);
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

// Launch arguments are read in place in the kernel-argument segment (the by-value
// parameter is the segment's first field and is not read directly), so each field is
// loaded where it is used instead of being held in scalar registers from the start.
template <typename T>
__device__ __forceinline__ const T &kernel_args() {
    return *(const T *)__builtin_amdgcn_kernarg_segment_ptr();
}

template <int kOcc, int kVar, bool kUseHiz, bool kUnderwater>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kOcc))) void k_march_tiles(TileArgs) {
    const TileArgs &a = kernel_args<TileArgs>();
    __shared__ uint64_t defmask[kUseHiz ? 128 : 1];
    __shared__ uint64_t selmask[(kVar & VAR_SPLIT) ? 128 : 1];
    const FrameK &f = a.f;
    const uint32_t ntiles = f.tiles[0] * f.tiles[1];
    if (blockIdx.x >= ntiles) return;
    const uint32_t tile = a.order[blockIdx.x];
    if (tile >= ntiles) return;
    // The order list starts with last frame's most expensive tiles: they are the
    // frame's critical path, so their waves get instruction-issue priority.
    if (blockIdx.x < kPrioTiles) __builtin_amdgcn_s_setprio(3);
    const uint64_t t0 = wall_clock64();
    const Spec s = a.s;
    march_tile_body<kUseHiz, kUnderwater, kVar>(a.p, f, s, tile, a.hiz, a.vis, a.shadow, a.cnt, defmask, selmask);
    if (lane_id() == 0) {
        uint64_t dt = wall_clock64() - t0;
        a.cost[tile] = dt > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dt;
    }
}

#define VX_TILES_INST1(O, V, H, U) template __global__ void k_march_tiles<O, V, H, U>(TileArgs);
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
VX_TILES_INST(6, 5)

// ---- Batched march: many views, one launch per view class ------------------------------
// A queue entry names (view, tile). The view's frame, spec and buffer offsets are read
// from its descriptor (wave-uniform, so scalar loads), and the tile runs the same body
// as the single-view kernel, so a pixel's key, shadow word and Hi-Z word are the same
// bits whichever way the views are launched. Views are laid out class by class (Hi-Z
// on/off x under water or not) and each class range is marched by a kernel specialised
// for it: one body per kernel keeps the register count at the single-view kernel's.
template <int kVar, bool kHiz, bool kUnderwater>
__device__ __forceinline__ void batch_tile(const BatchArgs &a, uint32_t entry, uint32_t n_views, uint64_t *defmask,
                                           uint64_t *selmask) {
    entry = (uint32_t)__builtin_amdgcn_readfirstlane((int)entry);  // uniform: descriptor reads are scalar
    const uint32_t v = entry >> ENTRY_VIEW_SHIFT, tile = entry & ENTRY_TILE_MASK;
    if (v >= n_views) return;
    // The descriptor is read through the constant address space with a uniform index:
    // scalar loads, and the frame fields live in SGPRs (the march indexes no frame array
    // with a run-time index, so the copy never needs addressable per-lane memory).
    const ViewK *vk = a.views + v;
    const FrameK &f = *(const FrameK *)as_const(&vk->f);
    const Spec s = ld_const(&vk->s);
    if (tile >= f.tiles[0] * f.tiles[1]) return;
    const_as<ViewK> *vc = as_const(vk);
    const Pools &p = *(const Pools *)as_const(&vk->p);
    const size_t px0 = (size_t)vc->px_off_lo | ((size_t)vc->px_off_hi << 32);
    const uint32_t toff = vc->tile_off;
    const uint32_t *bin_n = nullptr;
    const uint16_t *bin_l = nullptr;
    if (a.bin_cnt) {
        const uint32_t tx = f.tiles[0];
        const uint32_t bin = (tile / tx / BIN_TILES) * bins_x(tx) + (tile % tx) / BIN_TILES;
        bin_n = a.bin_cnt + vc->bin_off + bin;
        bin_l = a.bin_list + vc->binl_off + (size_t)bin * vc->bin_stride;
    }
    const uint64_t t0 = wall_clock64();
    march_tile_body<kHiz, kUnderwater, kVar>(p, f, s, tile, a.hiz + toff, a.vis + px0, a.shadow + px0, a.counters,
                                             defmask, selmask, bin_n, bin_l);
    if (lane_id() == 0) {
        uint64_t dt = wall_clock64() - t0;
        a.cost[toff + tile] = dt > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dt;
    }
}


// ---- Screen bins ------------------------------------------------------------------------
// ray_dir's pinhole model: pixel input (p, q) looks along fwd + right * sx(p) + up * sy(q),
// sx(p) = ((p + jitter_x) / W * 2 - 1) * tan * aspect, sy(q) = (1 - (q + jitter_y) / H * 2) * tan.
// A tile's side planes (tile_cull_setup) are the planes through the camera that hold the
// rays of its edges. A bin's side plane, moved BIN_MARGIN pixels outwards, has every
// tile's matching plane at least that far inside it, so a box wholly in front of the
// camera (positive along fwd in that model) and wholly outside the moved plane is at
// least BIN_MARGIN pixels outside each tile's frustum: no pixel ray of those tiles (at
// pixel centres) meets it, so no key changes, and tile_select's plane test rejects it
// unless that test's own f32 rounding is larger than the margin. A box is counted
// outside only when its largest signed distance is below -BIN_REL times a bound of the
// distances over the box, so that boxes that reach near the camera stay candidates.
constexpr float BIN_MARGIN = 2.0f, BIN_REL = 1e-3f;

__device__ __forceinline__ void cross3(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Largest value of n . (x - cam) over the box and a bound of |n . (x - cam)| over it.
__device__ __forceinline__ void box_plane_range(const float n[3], const float lo[3], const float hi[3], float &mx,
                                                float &bound) {
    mx = 0.0f;
    bound = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float a = n[i] * lo[i], b = n[i] * hi[i];
        mx += fmaxf(a, b);
        bound += fmaxf(fabsf(a), fabsf(b));
    }
}

__global__ __launch_bounds__(256) void k_bin_chunks(BatchArgs a, uint16_t *__restrict__ bin_list,
                                                    uint32_t *__restrict__ bin_cnt) {
    const uint32_t nb = a.hdr[BH_BINS], nv = a.hdr[BH_VIEWS];
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    uint32_t g = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    for (; g < nb; g += waves) {
        // The view of bin g: the last one whose first bin is <= g (bin_off grows with the slot).
        uint32_t lo = 0, hi = nv;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (as_const(a.views + mid)->bin_off <= g) lo = mid;
            else hi = mid;
        }
        const ViewK *vk = a.views + lo;
        const_as<ViewK> *vc = as_const(vk);
        const FrameK &f = *(const FrameK *)as_const(&vk->f);
        const Pools &p = *(const Pools *)as_const(&vk->p);
        const uint32_t b = g - vc->bin_off;
        const uint32_t nbx = bins_x(f.tiles[0]);
        const uint32_t bx = b % nbx, by = b / nbx;
        const uint32_t n = min(f.chunk_count, p.n_chunks);
        uint16_t *out = bin_list + vc->binl_off + (size_t)b * vc->bin_stride;
        // Views whose camera the model above does not cover well keep whole lists.
        const float th = f.tan_half_fov, ta = f.tan_half_fov * f.aspect;
        float rxu[3];
        cross3(f.right, f.up, rxu);
        const float det = dot3(f.fwd, rxu);
        const float nf = dot3(f.fwd, f.fwd), nr = dot3(f.right, f.right), nu = dot3(f.up, f.up);
        const bool usable = th > 0.0f && th < 8.0f && ta > 0.0f && ta < 8.0f && f.res[0] <= 16384u &&
                            f.res[1] <= 16384u && fabsf(f.jitter[0]) <= 4.0f && fabsf(f.jitter[1]) <= 4.0f &&
                            nf >= 0.25f && nf <= 4.0f && nr >= 0.25f && nr <= 4.0f && nu >= 0.25f && nu <= 4.0f &&
                            fabsf(det) >= 0.5f;
        float pl[5][3];  // left, right, top, bottom, front
        {
            const float W = (float)f.res[0], H = (float)f.res[1];
            constexpr float span = (float)(BIN_TILES * 8u);
            const float px0 = (float)(bx * BIN_TILES * 8u) - BIN_MARGIN, px1 = px0 + span + 2.0f * BIN_MARGIN;
            const float py0 = (float)(by * BIN_TILES * 8u) - BIN_MARGIN, py1 = py0 + span + 2.0f * BIN_MARGIN;
            auto sx = [&](float q) { return ((q + f.jitter[0]) / W * 2.0f - 1.0f) * ta; };
            auto sy = [&](float q) { return (1.0f - (q + f.jitter[1]) / H * 2.0f) * th; };
            const float cx = sx(px0 + 0.5f * span + BIN_MARGIN), cy = sy(py0 + 0.5f * span + BIN_MARGIN);
            float ce[3];
#pragma unroll
            for (int i = 0; i < 3; i++) ce[i] = f.fwd[i] + f.right[i] * cx + f.up[i] * cy;
            const float s4[4] = {sx(px0), sx(px1), sy(py0), sy(py1)};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float d[3];
#pragma unroll
                for (int i = 0; i < 3; i++) d[i] = f.fwd[i] + (k < 2 ? f.right[i] : f.up[i]) * s4[k];
                if (k < 2) cross3(d, f.up, pl[k]);
                else cross3(f.right, d, pl[k]);
                if (dot3(pl[k], ce) < 0.0f)
#pragma unroll
                    for (int i = 0; i < 3; i++) pl[k][i] = -pl[k][i];
            }
            const float sg = det < 0.0f ? -1.0f : 1.0f;
#pragma unroll
            for (int i = 0; i < 3; i++) pl[4][i] = rxu[i] * sg;
        }
        uint32_t count = 0;
        for (uint32_t w = 0; w < n; w += 64u) {
            const uint32_t ci = w + lane;
            bool keep = false;
            if (ci < n) {
                keep = true;
                if (usable && ci != f.camera_chunk) {
                    ChunkK c = ld_chunk(p, ci);
                    float lo3[3], hi3[3];
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        lo3[i] = c.bmin[i] - f.cam[i];
                        hi3[i] = c.bmax[i] - f.cam[i];
                    }
                    // in front: the smallest front distance over the box is positive
                    float neg[3] = {-pl[4][0], -pl[4][1], -pl[4][2]};
                    float mx, bound;
                    box_plane_range(neg, lo3, hi3, mx, bound);
                    const bool front = mx < -BIN_REL * bound;
                    bool out_side = false;
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        box_plane_range(pl[k], lo3, hi3, mx, bound);
                        out_side |= mx < -BIN_REL * bound;
                    }
                    keep = !(front && out_side);
                }
            }
            const uint64_t m = ballot(keep);
            if (keep) out[count + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)ci;
            count += (uint32_t)__popcll(m);
        }
        if (lane == 0) bin_cnt[g] = count;
    }
}

template <int kOcc, int kVar, bool kHiz, bool kUnderwater>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kOcc))) void k_batch_flat(BatchArgs) {
    const BatchArgs &a = kernel_args<BatchArgs>();
    __shared__ uint64_t defmask[kHiz ? 128 : 1];
    __shared__ uint64_t selmask[(kVar & VAR_SPLIT) ? 128 : 1];
    constexpr uint32_t cls = view_class(kHiz, kUnderwater);
    const uint32_t k = a.hdr[BH_CB + cls] + blockIdx.x;
    if (k >= a.hdr[BH_CB + cls + 1]) return;
    if (blockIdx.x < kPrioTiles) __builtin_amdgcn_s_setprio(3);
    batch_tile<kVar, kHiz, kUnderwater>(a, a.order[k], a.hdr[BH_VIEWS], defmask, selmask);
}

// Persistent form: the grid is sized to the resident wave slots; every wave takes
// `chunk` consecutive entries of its class range per atomic on the class's head word
// until the range is exhausted (the head only grows, so every wave reaches the exit).
// The range is sorted most expensive first, so small chunks keep the heavy head of the
// queue spread over many waves.
template <int kOcc, int kVar, bool kHiz, bool kUnderwater>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kOcc))) void k_batch_persistent(BatchArgs,
                                                                                                   uint32_t chunk) {
    const BatchArgs &a = kernel_args<BatchArgs>();
    __shared__ uint64_t defmask[kHiz ? 128 : 1];
    __shared__ uint64_t selmask[(kVar & VAR_SPLIT) ? 128 : 1];
    constexpr uint32_t cls = view_class(kHiz, kUnderwater);
    const uint32_t begin = a.hdr[BH_CB + cls], end = a.hdr[BH_CB + cls + 1];
    const uint32_t nv = a.hdr[BH_VIEWS];
    const uint32_t n = end > begin ? end - begin : 0u;
    uint32_t *head = a.counters + CNT_CTL + CTL_HEAD + cls * CTL_HEAD_STRIDE;
    const uint32_t grab = chunk ? chunk : 1u;
    for (;;) {
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(head, grab);
        base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);  // all lanes active: lane 0's value
        if (base >= n) break;
        if (base < kPrioTiles) __builtin_amdgcn_s_setprio(3);
        else __builtin_amdgcn_s_setprio(0);
        const uint32_t stop = min(base + grab, n);
        for (uint32_t k = base; k < stop; k++)
            batch_tile<kVar, kHiz, kUnderwater>(a, a.order[begin + k], nv, defmask, selmask);
    }
}

#define VX_BATCH_INST1(O, V, H, U)                                               \
    template __global__ void k_batch_flat<O, V, H, U>(BatchArgs);                \
    template __global__ void k_batch_persistent<O, V, H, U>(BatchArgs, uint32_t);
#define VX_BATCH_INST(O, V)              \
    VX_BATCH_INST1(O, V, true, false)   \
    VX_BATCH_INST1(O, V, false, false)  \
    VX_BATCH_INST1(O, V, true, true)    \
    VX_BATCH_INST1(O, V, false, true)
VX_BATCH_INST(4, 4)
VX_BATCH_INST(4, 5)
VX_BATCH_INST(6, 4)
VX_BATCH_INST(6, 5)

bool batch_form_ok(uint32_t mode, uint32_t occ, uint32_t var) {
    if (mode != VX_BATCH_FLAT && mode != VX_BATCH_PERSISTENT) return false;
    const uint32_t v = var & ~(uint32_t)VAR_BINS;  // bins: a launch argument, not a kernel form
    return (occ == 4 || occ == 6) && (v == 4 || v == 5);
}

template <int O, int V, bool H, bool U>
static void launch_class(uint32_t mode, uint32_t chunk, uint32_t wps, int n_cu, uint32_t n, const BatchArgs &a,
                         hipStream_t st) {
    if (mode == VX_BATCH_FLAT) k_batch_flat<O, V, H, U><<<n, 64, 0, st>>>(a);
    else k_batch_persistent<O, V, H, U><<<(unsigned)n_cu * 4u * wps, 64, 0, st>>>(a, chunk);
}

template <int O, int V>
static void launch_classes_ov(uint32_t mode, uint32_t chunk, uint32_t wps, int n_cu, const uint32_t cb[N_CLASSES + 1],
                              const BatchArgs &a, hipStream_t st) {
    for (uint32_t c = 0; c < N_CLASSES; c++) {
        const uint32_t n = cb[c + 1] - cb[c];
        if (!n) continue;
        switch (c) {
            case 0: launch_class<O, V, true, false>(mode, chunk, wps, n_cu, n, a, st); break;
            case 1: launch_class<O, V, false, false>(mode, chunk, wps, n_cu, n, a, st); break;
            case 2: launch_class<O, V, true, true>(mode, chunk, wps, n_cu, n, a, st); break;
            default: launch_class<O, V, false, true>(mode, chunk, wps, n_cu, n, a, st); break;
        }
    }
}

void launch_batch_classes(uint32_t mode, uint32_t occ, uint32_t var, uint32_t chunk, uint32_t wps, int n_cu,
                          const uint32_t cb[N_CLASSES + 1], const BatchArgs &a, hipStream_t st) {
    switch (occ * 16 + (var & ~(uint32_t)VAR_BINS)) {
        case 4 * 16 + 5: launch_classes_ov<4, 5>(mode, chunk, wps, n_cu, cb, a, st); break;
        case 6 * 16 + 4: launch_classes_ov<6, 4>(mode, chunk, wps, n_cu, cb, a, st); break;
        case 6 * 16 + 5: launch_classes_ov<6, 5>(mode, chunk, wps, n_cu, cb, a, st); break;
        default: launch_classes_ov<4, 4>(mode, chunk, wps, n_cu, cb, a, st); break;
    }
}

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

// Order bin of global tile i: its class (from the class boundaries, classes are laid out
// one after another) and its cost bucket, most expensive first inside the class.
struct ClassBounds {
    uint32_t b1, b2, b3;
    __device__ explicit ClassBounds(const uint32_t *cb)
        : b1(cb ? cb[1] : 0xFFFFFFFFu), b2(cb ? cb[2] : 0xFFFFFFFFu), b3(cb ? cb[3] : 0xFFFFFFFFu) {}
    __device__ uint32_t bin(uint32_t i, uint32_t cost) const {
        const uint32_t cls = (uint32_t)(i >= b1) + (uint32_t)(i >= b2) + (uint32_t)(i >= b3);
        return cls * COST_BUCKETS + (COST_BUCKETS - 1u - cost_bucket(cost));
    }
};

__global__ __launch_bounds__(256) void k_cost_hist(const uint32_t *__restrict__ cost, const uint32_t *__restrict__ n_dev,
                                                   uint32_t n, const uint32_t *__restrict__ cb,
                                                   uint32_t *__restrict__ ctl) {
    __shared__ uint32_t hist[ORDER_BINS];
    if (n_dev) n = *n_dev;
    const ClassBounds bounds(cb);
    const uint32_t t = threadIdx.x;
    for (uint32_t k = t; k < ORDER_BINS; k += blockDim.x) hist[k] = 0;
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + t; i < n; i += stride) atomicAdd(&hist[bounds.bin(i, cost[i])], 1u);
    __syncthreads();
    for (uint32_t k = t; k < ORDER_BINS; k += blockDim.x)
        if (hist[k]) atomicAdd(&ctl[CTL_HIST + k], hist[k]);
}

__global__ __launch_bounds__(256) void k_cost_scatter(const uint32_t *__restrict__ cost,
                                                      const uint32_t *__restrict__ n_dev, uint32_t n,
                                                      const uint32_t *__restrict__ cb, uint32_t *__restrict__ ctl,
                                                      const uint32_t *__restrict__ tile_off,
                                                      const uint32_t *__restrict__ n_views_dev,
                                                      uint32_t *__restrict__ order) {
    static_assert(ORDER_BINS == 128, "the bin scan below gives each lane of one wave two bins");
    __shared__ uint32_t cnt[ORDER_BINS];
    __shared__ uint32_t base[ORDER_BINS];
    __shared__ uint32_t toff[BATCH_MAX_VIEWS + 1];
    if (n_dev) n = *n_dev;
    const ClassBounds bounds(cb);
    const uint32_t nv = n_views_dev ? min(*n_views_dev, BATCH_MAX_VIEWS) : 0u;
    const uint32_t t = threadIdx.x;
    for (uint32_t k = t; k < ORDER_BINS; k += blockDim.x) cnt[k] = 0;
    if (nv)
        for (uint32_t v = t; v <= nv; v += blockDim.x) toff[v] = tile_off[v];
    // Bin starts: exclusive prefix sum of the grid-wide histogram, by the first wave (two
    // bins per lane, then a shuffle scan over the lanes' pair sums).
    if (t < 64u) {
        const uint32_t h0 = ctl[CTL_HIST + 2u * t], h1 = ctl[CTL_HIST + 2u * t + 1u];
        uint32_t s = h0 + h1;
        for (uint32_t d = 1; d < 64u; d <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)s, d);
            if (t >= d) s += o;
        }
        base[2u * t] = s - h0 - h1;
        base[2u * t + 1u] = s - h1;
    }
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + t; i < n; i += stride) atomicAdd(&cnt[bounds.bin(i, cost[i])], 1u);
    __syncthreads();
    for (uint32_t k = t; k < ORDER_BINS; k += blockDim.x) {
        const uint32_t c = cnt[k];
        base[k] = c ? base[k] + atomicAdd(&ctl[CTL_CURSOR + k], c) : 0u;
        cnt[k] = 0;
    }
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + t; i < n; i += stride) {
        const uint32_t k = bounds.bin(i, cost[i]);
        order[base[k] + atomicAdd(&cnt[k], 1u)] = tile_entry(i, toff, nv);
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
