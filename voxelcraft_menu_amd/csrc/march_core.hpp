// march_core.hpp -- device-side voxel DDA for gfx950.
//
// The arithmetic reproduces the reference WGSL bit for bit: f32, IEEE division and
// square root (this file is compiled with -ffp-contract=off and correctly rounded
// div/sqrt), WGSL operation order, minNum/maxNum for min/max. Each function cites
// the WGSL it computes. Pool reads are bounds-guarded: an index outside a pool reads
// zeros, so malformed input cannot take the GPU outside its buffers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vxmarch.h"

namespace vxd {

constexpr uint32_t FULL_BIT = 0x80000000u;
constexpr uint32_t PTR_MASK = 0x7FFFFFFFu;
constexpr uint32_t UNIFORM_BIT = 0x80000000u;
constexpr uint32_t NO_CHUNK = 0xFFFFFFFFu;
constexpr uint32_t MAX_ITERS = 512u;
constexpr uint32_t WATER_ID = 13u, GLASS_ID = 10u, GLOWSTONE_ID = 11u, LEAVES_ID = 7u, PINE_LEAVES_ID = 16u;
constexpr uint32_t ATTR_UNIFORM = 1u, ATTR_HAS_WATER = 2u, ATTR_HAS_CUTOUT = 8u, ATTR_HAS_GLASS = 16u,
                   ATTR_HAS_EMITTER = 32u;
constexpr uint32_t NORMAL_CROSS_A = 6u, NORMAL_CROSS_B = 7u;

// Pipeline specialisation bits (the WGSL override constants the march reads).
constexpr uint32_t SPEC_CUTOUT = 1u, SPEC_GLASS = 2u, SPEC_FULL_MARCH = 4u, SPEC_EMITTER = 8u,
                   SPEC_WATER_REPAIR = 16u, SPEC_GLASS_REFLECT = 32u,
                   // the water-look ripples (spec_hi), whose normals are not restated
                   SPEC_WATER_LOOK = 64u,
                   // the wave field (SPEC_WAVES, SPEC_WAVE_ANISO, SPEC_WAVE_SHOAL, SPEC_WAVE_FILL)
                   SPEC_WAVES = 128u, SPEC_WAVE_ANISO = 256u, SPEC_WAVE_SHOAL = 512u, SPEC_WAVE_FILL = 1024u,
                   SPEC_SNELL_BEND = 2048u;

struct Spec {
    uint32_t bits;
    float leaf_fill;
};

struct Pools {
    const uint4 *__restrict__ inners;
    const uint2 *__restrict__ leaves;
    const uint32_t *__restrict__ bricks;
    const vx_gpu_chunk *__restrict__ chunks;
    const uint32_t *__restrict__ grid;
    uint32_t n_inners, n_leaves, n_bricks, n_chunks;
};

// Frame fields read by the passes, packed for kernel arguments.
struct FrameK {
    float cam[3], fwd[3], right[3], up[3];
    float tan_half_fov, aspect, far_, daylight;
    float sun[3], shadow_dist, sea_level;
    float inv_sun[3];  // 1.0 / sun per component (IEEE division on the host): the sun ray's inv
    float jitter[2];
    uint32_t res[2], tiles[2];
    uint32_t chunk_count, camera_chunk, flags, max_pairs, frame_index;
    int32_t grid_min[3];
    uint32_t grid_dim[3];
    float time, wave_amp, wave_scale, wave_speed, wave_reflect_slope;  // the wave field's inputs
};

// Pool reads go through buffer resources whose size is the pool's length: an index past
// the end reads zeros (the restatement's reading of robust buffer access) by the
// hardware's range check instead of a compare and branch per fetch. Pools stay below
// 4 GiB - 16 bytes (vx_pool_sync refuses more), and the index is clamped first so that
// index * element size cannot wrap back into range.
constexpr uint32_t BUF_DWORD3 = 0x00020000u;  // gfx9 buffer descriptor word 3: 32-bit data format
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pool_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, (int)BUF_DWORD3);
}
__device__ __forceinline__ uint4 ld_inner(const Pools &p, uint32_t i) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(pool_rsrc(p.inners, p.n_inners * 16u),
                                                         min(i, 0x0FFFFFFFu) << 4, 0, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ uint2 ld_leaf(const Pools &p, uint32_t i) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(pool_rsrc(p.leaves, p.n_leaves * 8u),
                                                        min(i, 0x1FFFFFFFu) << 3, 0, 0);
    return make_uint2(v[0], v[1]);
}
__device__ __forceinline__ uint32_t ld_brick(const Pools &p, uint32_t i) {
    return __builtin_amdgcn_raw_buffer_load_b32(pool_rsrc(p.bricks, p.n_bricks * 4u), min(i, 0x3FFFFFFFu) << 2,
                                                0, 0);
}

// The chunk fields the march reads from a GpuChunk record (96 B).
struct ChunkK {
    float bmin[3], vs, bmax[3], origin[3], fade;
    uint32_t root[4];
    uint32_t attr_base, attr_flags;
    uint32_t dry[2];
};

// Read-only data of a launch (chunk records, view descriptors) is read through the
// constant address space: with a wave-uniform index the loads become scalar loads into
// SGPRs instead of one vector load per lane.
template <typename T>
using const_as = const __attribute__((address_space(4))) T;
template <typename T>
__device__ __forceinline__ const_as<T> *as_const(const T *p) {
    return (const_as<T> *)p;
}
// Copy of a trivially copyable record, read word by word through the constant space.
template <typename T>
__device__ __forceinline__ T ld_const(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "record of whole words");
    T out;
    const_as<uint32_t> *src = as_const(reinterpret_cast<const uint32_t *>(p));
    uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4; i++) dst[i] = src[i];
    return out;
}

__device__ __forceinline__ ChunkK ld_chunk(const Pools &p, uint32_t ci) {
    ChunkK c;
    if (ci >= p.n_chunks) ci = 0;  // callers guard; keeps the read inside the buffer
    const_as<uint32_t> *w = as_const(reinterpret_cast<const uint32_t *>(p.chunks + ci));
    c.bmin[0] = __uint_as_float(w[0]); c.bmin[1] = __uint_as_float(w[1]); c.bmin[2] = __uint_as_float(w[2]);
    c.vs = __uint_as_float(w[3]);
    c.bmax[0] = __uint_as_float(w[4]); c.bmax[1] = __uint_as_float(w[5]); c.bmax[2] = __uint_as_float(w[6]);
    c.origin[0] = __uint_as_float(w[8]); c.origin[1] = __uint_as_float(w[9]); c.origin[2] = __uint_as_float(w[10]);
    c.fade = __uint_as_float(w[11]);
    c.root[0] = w[12]; c.root[1] = w[13]; c.root[2] = w[14]; c.root[3] = w[15];
    c.attr_base = w[16]; c.attr_flags = w[17];
    c.dry[0] = w[20]; c.dry[1] = w[21];
    return c;
}

// x / vs with the division skipped when vs is exactly 1.0 (x / 1.0f == x in IEEE f32,
// so the result is bit-identical); LOD-0 chunks always have vs == 1.
__device__ __forceinline__ float div_vs(float x, float vs) { return vs == 1.0f ? x : x / vs; }
__device__ __forceinline__ float mul_vs(float x, float vs) { return vs == 1.0f ? x : x * vs; }

// mask_has / mask_below / cell_bit (common.wgsl:299-318) on a u64 mask.
__device__ __forceinline__ bool mask_has(uint64_t m, uint32_t bit) { return (m >> bit) & 1ull; }
__device__ __forceinline__ uint32_t mask_below(uint64_t m, uint32_t bit) {
    return (uint32_t)__popcll(m & ((1ull << bit) - 1ull));  // bit < 64 on every call
}
__device__ __forceinline__ uint32_t cell_bit(uint32_t x, uint32_t y, uint32_t z) { return x | (y << 2) | (z << 4); }
__device__ __forceinline__ uint64_t mk64(uint32_t lo, uint32_t hi) { return (uint64_t)lo | ((uint64_t)hi << 32); }

// leaf_block (common.wgsl:320-334)
__device__ __forceinline__ uint32_t leaf_block(const Pools &p, uint32_t entry, uint32_t vx, uint32_t vy, uint32_t vz) {
    if (entry & UNIFORM_BIT) return entry & 0xFFFFu;
    uint32_t hdr = ld_brick(p, entry);
    uint32_t n = hdr & 0xFFu;
    uint32_t bits = (hdr >> 8) & 0xFFu;
    uint32_t pal_words = (n + 1u) / 2u;
    uint32_t vb = cell_bit(vx & 3u, vy & 3u, vz & 3u);
    uint32_t bitpos = vb * bits;
    uint32_t w = ld_brick(p, entry + 1u + pal_words + (bitpos >> 5));
    uint32_t idx = (w >> (bitpos & 31u)) & ((1u << (bits & 31u)) - 1u);
    uint32_t pw = ld_brick(p, entry + 1u + idx / 2u);
    return (pw >> ((idx & 1u) * 16u)) & 0xFFFFu;
}

// leaf_attr + block_at (common.wgsl:336-356)
__device__ __forceinline__ uint32_t block_at(const Pools &p, const ChunkK &c, uint32_t vx, uint32_t vy, uint32_t vz) {
    if (c.attr_flags & ATTR_UNIFORM) return c.attr_base;
    uint64_t root_mask = mk64(c.root[0], c.root[1]);
    uint32_t l1b = cell_bit(vx >> 4, vy >> 4, vz >> 4);
    uint4 l1;
    if (c.root[2] & FULL_BIT) l1 = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, FULL_BIT, l1b * 64u);
    else l1 = ld_inner(p, (c.root[2] & PTR_MASK) + mask_below(root_mask, l1b));
    uint32_t lfb = cell_bit((vx >> 2) & 3u, (vy >> 2) & 3u, (vz >> 2) & 3u);
    uint32_t entry = ld_brick(p, c.attr_base + (l1.w & 0x0FFFFFFFu) + mask_below(mk64(l1.x, l1.y), lfb));
    return leaf_block(p, entry, vx, vy, vz);
}

// voxel_occupied (common.wgsl:912-940)
__device__ inline bool voxel_occupied(const Pools &p, const ChunkK &c, int32_t x, int32_t y, int32_t z) {
    if (x < 0 || y < 0 || z < 0 || x > 63 || y > 63 || z > 63) return false;
    uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
    uint64_t root_mask = mk64(c.root[0], c.root[1]);
    uint32_t l1b = cell_bit(ux >> 4, uy >> 4, uz >> 4);
    if (!mask_has(root_mask, l1b)) return false;
    if (c.root[2] & FULL_BIT) return true;
    uint4 l1 = ld_inner(p, (c.root[2] & PTR_MASK) + mask_below(root_mask, l1b));
    if (l1.z & FULL_BIT) return true;
    uint64_t l1m = mk64(l1.x, l1.y);
    uint32_t lfb = cell_bit((ux >> 2) & 3u, (uy >> 2) & 3u, (uz >> 2) & 3u);
    if (!mask_has(l1m, lfb)) return false;
    uint2 lf = ld_leaf(p, (l1.z & PTR_MASK) + mask_below(l1m, lfb));
    return mask_has(mk64(lf.x, lf.y), cell_bit(ux & 3u, uy & 3u, uz & 3u));
}

__device__ __forceinline__ uint32_t grid_index(const FrameK &f, int32_t cx, int32_t cy, int32_t cz, bool &ok) {
    int32_t rx = cx - f.grid_min[0], ry = cy - f.grid_min[1], rz = cz - f.grid_min[2];
    ok = rx >= 0 && ry >= 0 && rz >= 0 && rx < (int32_t)f.grid_dim[0] && ry < (int32_t)f.grid_dim[1] &&
         rz < (int32_t)f.grid_dim[2];
    return (uint32_t)rx + (uint32_t)ry * f.grid_dim[0] + (uint32_t)rz * f.grid_dim[0] * f.grid_dim[1];
}

__device__ __forceinline__ uint32_t ld_grid(const Pools &p, const FrameK &f, uint32_t idx) {
    return idx < f.grid_dim[0] * f.grid_dim[1] * f.grid_dim[2] && idx < VX_GRID_LEN ? p.grid[idx] : NO_CHUNK;
}

// world_block_at (common.wgsl:266-277)
__device__ inline uint32_t world_block_at(const Pools &p, const FrameK &f, float px, float py, float pz) {
    bool ok;
    uint32_t gi = grid_index(f, (int32_t)floorf(px / 64.0f), (int32_t)floorf(py / 64.0f), (int32_t)floorf(pz / 64.0f), ok);
    if (!ok) return 0u;
    uint32_t ci = ld_grid(p, f, gi);
    if (ci == NO_CHUNK || ci >= p.n_chunks) return 0u;
    ChunkK c = ld_chunk(p, ci);
    int32_t vx = (int32_t)floorf(div_vs(px - c.origin[0], c.vs));
    int32_t vy = (int32_t)floorf(div_vs(py - c.origin[1], c.vs));
    int32_t vz = (int32_t)floorf(div_vs(pz - c.origin[2], c.vs));
    if (vx < 0 || vy < 0 || vz < 0 || vx >= 64 || vy >= 64 || vz >= 64) return 0u;
    if (!voxel_occupied(p, c, vx, vy, vz)) return 0u;
    return block_at(p, c, (uint32_t)vx, (uint32_t)vy, (uint32_t)vz);
}

// world_block_at for a point beside the cell a march is on (the water-interface probe of
// march_chunk, common.wgsl:800-813): the grid cell is looked up exactly as world_block_at
// does; when it names the chunk being marched (ci), the point's voxel is read from that
// chunk with the march's L1 node, leaf and attribute word reused where the voxel shares
// them (same L1 cell, same leaf), instead of reloading the record and the node chain.
// The nodes are the same words world_block_at would load, so the block id is the same.
template <bool kUnit>
__device__ inline uint32_t world_block_near(const Pools &p, const FrameK &f, const ChunkK &c, uint32_t ci, float px,
                                            float py, float pz, uint32_t l1_cached, const uint4 &l1,
                                            uint32_t leaf_cached, uint64_t leaf, uint32_t attr) {
    bool ok;
    const uint32_t gi =
        grid_index(f, (int32_t)floorf(px / 64.0f), (int32_t)floorf(py / 64.0f), (int32_t)floorf(pz / 64.0f), ok);
    if (!ok) return 0u;
    const uint32_t ci2 = ld_grid(p, f, gi);
    if (ci2 == NO_CHUNK || ci2 >= p.n_chunks) return 0u;
    if (ci2 != ci) return world_block_at(p, f, px, py, pz);
    auto dvs = [&](float x) { return kUnit ? x : x / c.vs; };
    const int32_t vx = (int32_t)floorf(dvs(px - c.origin[0]));
    const int32_t vy = (int32_t)floorf(dvs(py - c.origin[1]));
    const int32_t vz = (int32_t)floorf(dvs(pz - c.origin[2]));
    if (vx < 0 || vy < 0 || vz < 0 || vx >= 64 || vy >= 64 || vz >= 64) return 0u;
    const uint32_t ux = (uint32_t)vx, uy = (uint32_t)vy, uz = (uint32_t)vz;
    const uint64_t root_mask = mk64(c.root[0], c.root[1]);
    const uint32_t l1b = cell_bit(ux >> 4, uy >> 4, uz >> 4);
    if (!mask_has(root_mask, l1b)) return 0u;  // voxel_occupied
    const bool root_full = (c.root[2] & FULL_BIT) != 0u;
    const bool same_l1 = l1b == l1_cached;
    uint4 n1;
    if (same_l1) n1 = l1;
    else if (root_full) n1 = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, FULL_BIT, l1b * 64u);
    else n1 = ld_inner(p, (c.root[2] & PTR_MASK) + mask_below(root_mask, l1b));
    const uint64_t l1m = mk64(n1.x, n1.y);
    const uint32_t lfb = cell_bit((ux >> 2) & 3u, (uy >> 2) & 3u, (uz >> 2) & 3u);
    const bool same_leaf = same_l1 && lfb == leaf_cached;
    if (!root_full && !(n1.z & FULL_BIT)) {
        if (!mask_has(l1m, lfb)) return 0u;
        uint64_t lm = leaf;
        if (!same_leaf) {
            const uint2 lf = ld_leaf(p, (n1.z & PTR_MASK) + mask_below(l1m, lfb));
            lm = mk64(lf.x, lf.y);
        }
        if (!mask_has(lm, cell_bit(ux & 3u, uy & 3u, uz & 3u))) return 0u;
    }
    if (c.attr_flags & ATTR_UNIFORM) return c.attr_base;  // block_at
    const uint32_t entry = same_leaf ? attr : ld_brick(p, c.attr_base + (n1.w & 0x0FFFFFFFu) + mask_below(l1m, lfb));
    return leaf_block(p, entry, ux, uy, uz);
}

// ray_dir (common.wgsl:358-368): normalize = v / sqrt(dot(v, v)).
__device__ __forceinline__ void ray_dir(const FrameK &f, float px, float py, float &dx, float &dy, float &dz) {
    float ndc_x = ((px + f.jitter[0]) / (float)f.res[0]) * 2.0f - 1.0f;
    float ndc_y = 1.0f - ((py + f.jitter[1]) / (float)f.res[1]) * 2.0f;
    float sx = ndc_x * f.tan_half_fov * f.aspect;
    float sy = ndc_y * f.tan_half_fov;
    float x = f.fwd[0] + f.right[0] * sx + f.up[0] * sy;
    float y = f.fwd[1] + f.right[1] * sx + f.up[1] * sy;
    float z = f.fwd[2] + f.right[2] * sx + f.up[2] * sy;
    float len = sqrtf(x * x + y * y + z * z);
    x = x / len;
    y = y / len;
    z = z / len;
    auto guard = [](float d) {
        if (fabsf(d) < 1e-6f) {
            float s = d + 1e-9f;
            d = 1e-6f * (s > 0.0f ? 1.0f : (s < 0.0f ? -1.0f : 0.0f));
        }
        return d;
    };
    dx = guard(x);
    dy = guard(y);
    dz = guard(z);
}

// ray_skips with dither (common.wgsl:370-385)
__device__ __forceinline__ bool ray_skips(const FrameK &f, float fade, uint32_t px, uint32_t py) {
    if (fade >= 1.0f) return false;
    float p0 = (float)px + 5.588238f * (float)f.frame_index;
    float p1 = (float)py;
    float a = p0 * 0.06711056f + p1 * 0.00583715f;
    a = a - floorf(a);
    float d = 52.9829189f * a;
    d = d - floorf(d);
    if (fade >= 0.0f) return d >= fade;
    return d < -fade;
}

// hash_u32 (common.wgsl:1422-1430)
__device__ __forceinline__ uint32_t hash_u32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

// leaf_solid (common.wgsl:514-518)
__device__ __forceinline__ bool leaf_solid(float fill, int32_t wx, int32_t wy, int32_t wz, uint32_t mx, uint32_t my,
                                           uint32_t mz) {
    uint32_t qx = (uint32_t)wx * 4u + mx, qy = (uint32_t)wy * 4u + my, qz = (uint32_t)wz * 4u + mz;
    uint32_t h = hash_u32((qx * 73856093u) ^ (qy * 19349663u) ^ (qz * 83492791u));
    return (float)(h >> 8) * (1.0f / 16777216.0f) < fill;
}

__device__ __forceinline__ int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// micro_entry (common.wgsl:520-528); returns the three 2-bit coordinates packed x | y<<2 | z<<4.
__device__ __forceinline__ uint32_t micro_entry(uint32_t ux, uint32_t uy, uint32_t uz, const float ro[3],
                                                const float rd[3], float t, uint32_t axis, const bool pd[3],
                                                bool snap) {
    int32_t m[3];
    uint32_t u[3] = {ux, uy, uz};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float q = (ro[i] + rd[i] * t - (float)u[i]) * 4.0f;
        m[i] = clampi((int32_t)floorf(q), 0, 3);
    }
    if (snap) {
        int32_t s = pd[axis] ? 0 : 3;
        if (axis == 0) m[0] = s;
        else if (axis == 1) m[1] = s;
        else m[2] = s;
    }
    return (uint32_t)m[0] | ((uint32_t)m[1] << 2) | ((uint32_t)m[2] << 4);
}

struct MicroHit {
    bool hit;
    float t;
    uint32_t normal;
    uint32_t micro;  // packed x | y<<2 | z<<4
};

// cutout_march (common.wgsl:537-577)
__device__ inline MicroHit cutout_march(float fill, int32_t wx, int32_t wy, int32_t wz, uint32_t ux, uint32_t uy,
                                        uint32_t uz, const float ro[3], const float rd[3], const float inv[3],
                                        float t_in, uint32_t axis_in, const bool pd[3], bool snap) {
    MicroHit r{false, t_in, 0u, 0u};
    const float base[3] = {(float)ux, (float)uy, (float)uz};
    // inv is 1/rd computed exactly as cutout_march's own `let inv = 1.0 / rd`.
    uint32_t me = micro_entry(ux, uy, uz, ro, rd, t_in, axis_in, pd, snap);
    int32_t m[3] = {(int32_t)(me & 3u), (int32_t)((me >> 2) & 3u), (int32_t)((me >> 4) & 3u)};
    float t = t_in;
    uint32_t axis = axis_in;
    for (uint32_t i = 0; i < 12u; i++) {
        if (m[0] < 0 || m[1] < 0 || m[2] < 0 || m[0] > 3 || m[1] > 3 || m[2] > 3) return r;
        if (leaf_solid(fill, wx, wy, wz, (uint32_t)m[0], (uint32_t)m[1], (uint32_t)m[2])) {
            r.hit = true;
            r.t = t;
            r.normal = axis * 2u + (pd[axis] ? 0u : 1u);
            r.micro = (uint32_t)m[0] | ((uint32_t)m[1] << 2) | ((uint32_t)m[2] << 4);
            return r;
        }
        float tn[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float bound = (float)(pd[k] ? m[k] + 1 : m[k]);
            tn[k] = (base[k] + bound * 0.25f - ro[k]) * inv[k];
        }
        if (tn[0] <= tn[1] && tn[0] <= tn[2]) { t = tn[0]; axis = 0; }
        else if (tn[1] <= tn[2]) { t = tn[1]; axis = 1; }
        else { t = tn[2]; axis = 2; }
        int32_t step = pd[axis] ? 1 : -1;
        if (axis == 0) m[0] += step;
        else if (axis == 1) m[1] += step;
        else m[2] += step;
    }
    return r;
}

struct Hit {
    bool hit;
    float t;
    uint32_t normal;
    uint32_t voxel;  // x | y<<6 | z<<12
    uint32_t micro;  // x | y<<2 | z<<4
    uint32_t iters;
};

// march_chunk (common.wgsl:579-910). Foliage tufts are not handled here: the
// context refuses SPEC_FOLIAGE, so foliage_aware is false on every call.
// inv is the ray's 1.0 / rd (the same value march_chunk's own `let inv = 1.0 / rd`
// computes for every chunk; callers compute it once per ray). kUnit: the chunk's voxel
// size is exactly 1.0f, so the divisions and products by it are the identity.
// The ray kind (skip_water, skip_glass, sun_ray) is a compile-time parameter: primary
// rays use <false,false,false> (or <true,false,false> under water), sun rays
// <true,true,true>, so each kernel carries only the branches its rays can take.
template <bool skip_water, bool skip_glass, bool sun_ray, bool kUnit>
__device__ __forceinline__ Hit march_chunk_vs(const Pools &p, const FrameK &f, const Spec &s, const ChunkK &c,
                                              uint32_t ci, const float ro_w[3], const float rd[3], const float inv[3],
                                              float t_limit) {
    Hit h{false, 0.0f, 0u, 0u, 0u, 0u};
    const float vs = c.vs;
    // x / vs and x * vs; with kUnit (vs == 1.0f) both are x exactly.
    auto dvs = [&](float x, float v) { return kUnit ? x : x / v; };
    auto mvs = [&](float x, float v) { return kUnit ? x : x * v; };
    float ro[3], bmin[3], bmax[3], tmin3[3], tmax3[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        ro[i] = dvs(ro_w[i] - c.origin[i], vs);
        bmin[i] = dvs(c.bmin[i] - c.origin[i], vs);
        bmax[i] = dvs(c.bmax[i] - c.origin[i], vs);
        float t0 = (bmin[i] - ro[i]) * inv[i];
        float t1 = (bmax[i] - ro[i]) * inv[i];
        tmin3[i] = fminf(t0, t1);
        tmax3[i] = fmaxf(t0, t1);
    }
    const float t_enter = fmaxf(fmaxf(tmin3[0], tmin3[1]), tmin3[2]);
    const float t_exit = fminf(fminf(tmax3[0], tmax3[1]), tmax3[2]);
    const float t_end = fminf(t_exit, dvs(t_limit, vs));
    float t = fmaxf(t_enter, 0.0f);
    if (t >= t_end) return h;
    uint32_t axis;
    if (tmin3[0] >= tmin3[1] && tmin3[0] >= tmin3[2]) axis = 0;
    else if (tmin3[1] >= tmin3[2]) axis = 1;
    else axis = 2;
    const bool pd[3] = {rd[0] > 0.0f, rd[1] > 0.0f, rd[2] > 0.0f};
    // All ones along + axes. Kept as per-lane integers in VGPRs (the empty asm hides that
    // they are booleans): as lane masks they compete for the scalar registers.
    int32_t pdm[3] = {pd[0] ? -1 : 0, pd[1] ? -1 : 0, pd[2] ? -1 : 0};
#pragma unroll
    for (int i = 0; i < 3; i++) asm volatile("" : "+v"(pdm[i]));
    int32_t cell[3];
#pragma unroll
    for (int i = 0; i < 3; i++) cell[i] = (int32_t)floorf(ro[i] + rd[i] * t);
    if (t_enter > 0.0f) {
        int32_t v = pd[axis] ? (int32_t)floorf(bmin[axis]) : (int32_t)ceilf(bmax[axis]) - 1;
        if (axis == 0) cell[0] = v;
        else if (axis == 1) cell[1] = v;
        else cell[2] = v;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) cell[i] = clampi(cell[i], 0, 63);
    const uint64_t root_mask = mk64(c.root[0], c.root[1]);
    const uint32_t root_ptr = c.root[2] & PTR_MASK;
    const bool attr_uniform = (c.attr_flags & ATTR_UNIFORM) != 0u;

    bool water_aware = skip_water && (c.attr_flags & ATTR_HAS_WATER) != 0u;
    if (water_aware && attr_uniform) {
        if (c.attr_base == WATER_ID) return h;
        water_aware = false;
    }
    bool cutout_aware = false;
    if (s.bits & SPEC_CUTOUT) {
        cutout_aware = (c.attr_flags & ATTR_HAS_CUTOUT) != 0u && vs == 1.0f;
        if (cutout_aware && attr_uniform && c.attr_base != LEAVES_ID && c.attr_base != PINE_LEAVES_ID)
            cutout_aware = false;
    }
    bool glass_aware = false;
    if (s.bits & SPEC_GLASS) {
        glass_aware = skip_glass && (c.attr_flags & ATTR_HAS_GLASS) != 0u;
        if (glass_aware && attr_uniform) {
            if (c.attr_base == GLASS_ID) return h;
            glass_aware = false;
        }
    }
    bool noncaster_aware = false;
    if (s.bits & SPEC_EMITTER) noncaster_aware = sun_ray && (c.attr_flags & ATTR_HAS_EMITTER) != 0u;
    if (noncaster_aware && attr_uniform && c.attr_base == GLOWSTONE_ID) return h;
    const bool water_interface =
        (s.bits & SPEC_WATER_REPAIR) && !skip_water && (c.attr_flags & ATTR_HAS_WATER) != 0u;
    const bool attr_needed = (noncaster_aware && !attr_uniform) || (water_interface && !attr_uniform) ||
                             water_aware || glass_aware || (cutout_aware && !attr_uniform);
    const uint64_t trav_mask = water_aware ? mk64(c.dry[0], c.dry[1]) : root_mask;
    const bool root_full = (c.root[2] & FULL_BIT) != 0u;
    if (root_full && !(s.bits & SPEC_FULL_MARCH) && !noncaster_aware && !water_interface) {
        h.hit = true;
        h.t = mvs(t, vs);
        h.normal = axis * 2u + (pd[axis] ? 0u : 1u);
        h.voxel = (uint32_t)cell[0] | ((uint32_t)cell[1] << 6) | ((uint32_t)cell[2] << 12);
        h.micro = micro_entry((uint32_t)cell[0], (uint32_t)cell[1], (uint32_t)cell[2], ro, rd, t, axis, pd, t > 0.0f);
        return h;
    }
    if (trav_mask == 0ull) return h;

    uint32_t l1_cached = 0xFFFFFFFFu, leaf_cached = 0xFFFFFFFFu;
    uint4 l1 = make_uint4(0, 0, 0, 0);
    uint64_t leaf = 0;
    uint32_t attr = 0;
    uint32_t iters = 0;
    for (;;) {
        iters++;
        if (iters > MAX_ITERS) break;
        if ((uint32_t)cell[0] > 63u || (uint32_t)cell[1] > 63u || (uint32_t)cell[2] > 63u) break;
        const uint32_t ux = (uint32_t)cell[0], uy = (uint32_t)cell[1], uz = (uint32_t)cell[2];
        int32_t skip = 16;
        bool solid = false;
        const uint32_t l1b = cell_bit(ux >> 4, uy >> 4, uz >> 4);
        if (mask_has(trav_mask, l1b)) {
            if (l1b != l1_cached) {
                l1_cached = l1b;
                if (root_full) l1 = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, FULL_BIT, (l1b * 64u) | 0xF0000000u);
                else l1 = ld_inner(p, root_ptr + mask_below(root_mask, l1b));
                leaf_cached = 0xFFFFFFFFu;
            }
            bool line_skip = false;
            if (water_aware) {
                uint32_t wl = l1.w >> 28;
                float ly = (float)(uy & 15u);
                if (wl < 15u && ly > (float)wl) {
                    line_skip = rd[1] >= 0.0f;
                    if (!line_skip) {
                        float hh = fmaxf(fabsf(rd[0]), fabsf(rd[2]));
                        if (hh > 1e-6f && (ly - (float)wl) > 16.0f * (fabsf(rd[1]) / hh)) line_skip = true;
                    }
                }
            }
            if (line_skip) {
                skip = 16;
            } else if (l1.z & FULL_BIT) {
                solid = true;
                if (attr_needed) {
                    uint32_t lfb = cell_bit((ux >> 2) & 3u, (uy >> 2) & 3u, (uz >> 2) & 3u);
                    if (lfb != leaf_cached) {
                        leaf_cached = lfb;
                        attr = ld_brick(p, c.attr_base + (l1.w & 0x0FFFFFFFu) + lfb);
                    }
                }
            } else {
                const uint64_t l1m = mk64(l1.x, l1.y);
                const uint32_t lfb = cell_bit((ux >> 2) & 3u, (uy >> 2) & 3u, (uz >> 2) & 3u);
                if (mask_has(l1m, lfb)) {
                    if (lfb != leaf_cached) {
                        leaf_cached = lfb;
                        uint32_t li = mask_below(l1m, lfb);
                        uint2 lf = ld_leaf(p, (l1.z & PTR_MASK) + li);
                        leaf = mk64(lf.x, lf.y);
                        if (attr_needed) attr = ld_brick(p, c.attr_base + (l1.w & 0x0FFFFFFFu) + li);
                    }
                    const uint32_t vb = cell_bit(ux & 3u, uy & 3u, uz & 3u);
                    if (mask_has(leaf, vb)) {
                        solid = true;
                    } else {
                        // 2x2x2 window starting at the even corner: bits (vb & 42) + {0,1,4,5,16,17,20,21}.
                        uint32_t lo = (uint32_t)(leaf >> (vb & 42u));
                        skip = ((lo & 0x00330033u) == 0u) ? 2 : 1;
                    }
                } else {
                    skip = 4;
                }
            }
        } else {
            uint32_t rlo = (uint32_t)(trav_mask >> (l1b & 42u));
            skip = ((rlo & 0x00330033u) == 0u) ? 32 : 16;
        }
        if (solid && noncaster_aware) {
            uint32_t id = attr_uniform ? c.attr_base : leaf_block(p, attr, ux, uy, uz);
            if (id == GLOWSTONE_ID) { solid = false; skip = 1; }
        }
        if (solid && water_interface) {
            uint32_t id = attr_uniform ? c.attr_base : leaf_block(p, attr, ux, uy, uz);
            if (id == WATER_ID) {
                // normal_of(axis*2 + (pos ? 0 : 1)): unit vector along axis, -1 if the ray goes +.
                float o[3] = {0.0f, 0.0f, 0.0f};
                float sgn = pd[axis] ? -1.0f : 1.0f;
                if (axis == 0) o[0] = sgn;
                else if (axis == 1) o[1] = sgn;
                else o[2] = sgn;
                float tw = mvs(t, vs);
                float off = mvs(0.002f, vs);
                float b0 = ro_w[0] + rd[0] * tw, b1 = ro_w[1] + rd[1] * tw, b2 = ro_w[2] + rd[2] * tw;
                if (world_block_near<kUnit>(p, f, c, ci, b0 + o[0] * off, b1 + o[1] * off, b2 + o[2] * off, l1_cached,
                                            l1, leaf_cached, leaf, attr) == WATER_ID) {
                    solid = false;
                    skip = 1;
                }
            }
        }
        if (solid && water_aware) {
            if (leaf_block(p, attr, ux, uy, uz) == WATER_ID) {
                solid = false;
                skip = (attr & UNIFORM_BIT) ? 4 : 1;
            }
        }
        if (solid && glass_aware) {
            if (leaf_block(p, attr, ux, uy, uz) == GLASS_ID) {
                solid = false;
                skip = (attr & UNIFORM_BIT) ? 4 : 1;
            }
        }
        if (solid && cutout_aware) {
            uint32_t bid = attr_uniform ? c.attr_base : leaf_block(p, attr, ux, uy, uz);
            if (bid == LEAVES_ID || bid == PINE_LEAVES_ID) {
                int32_t wx = (int32_t)c.origin[0] + (int32_t)ux, wy = (int32_t)c.origin[1] + (int32_t)uy,
                        wz = (int32_t)c.origin[2] + (int32_t)uz;
                MicroHit q = cutout_march(s.leaf_fill, wx, wy, wz, ux, uy, uz, ro, rd, inv, t, axis, pd, t > 0.0f);
                if (q.hit) {
                    h.hit = true;
                    h.t = mvs(q.t, vs);
                    h.normal = q.normal;
                    h.voxel = ux | (uy << 6) | (uz << 12);
                    h.micro = q.micro;
                    h.iters = iters;
                    return h;
                }
                solid = false;
                skip = 1;
            }
        }
        if (solid) {
            h.hit = true;
            h.t = mvs(t, vs);
            h.normal = axis * 2u + (pd[axis] ? 0u : 1u);
            h.voxel = ux | (uy << 6) | (uz << 12);
            h.micro = micro_entry(ux, uy, uz, ro, rd, t, axis, pd, t > 0.0f);
            h.iters = iters;
            return h;
        }
        // Step to the exit face of the skip box (common.wgsl:888-907), written with
        // selects only: box = cell - (cell & (skip - 1)) = cell & -skip; the exit bound
        // is box + skip along + axes and box along - axes; the first axis wins ties
        // (x <= y <= z); the new cell is floor(ro + rd * t) with the stepped axis forced
        // to the next box, box + skip or box - 1.
        int32_t box[3];
        float tn[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            box[i] = cell[i] & -skip;
            tn[i] = ((float)(box[i] + (skip & pdm[i])) - ro[i]) * inv[i];
        }
        const bool s0 = (tn[0] <= tn[1]) & (tn[0] <= tn[2]);
        const bool s1 = !s0 & (tn[1] <= tn[2]);
        const bool s2 = !s0 & !s1;
        t = s0 ? tn[0] : (s1 ? tn[1] : tn[2]);
        if (t >= t_end) break;
        axis = s0 ? 0u : (s1 ? 1u : 2u);
        const bool sel[3] = {s0, s1, s2};
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int32_t next = box[i] + ((skip + 1) & pdm[i]) - 1;
            const int32_t fl = (int32_t)floorf(ro[i] + rd[i] * t);
            cell[i] = sel[i] ? next : fl;
        }
    }
    h.iters = iters;
    return h;
}

// march_chunk dispatches on the chunk's voxel size: LOD-0 chunks (vs == 1.0f) take the
// variant with no division or multiplication by vs, which is bit-identical because
// x / 1.0f == x * 1.0f == x in IEEE f32. Without the split the compiler evaluates the
// correctly rounded divisions of the general case and then selects x.
template <bool skip_water, bool skip_glass, bool sun_ray>
__device__ __forceinline__ Hit march_chunk(const Pools &p, const FrameK &f, const Spec &s, const ChunkK &c, uint32_t ci,
                                           const float ro_w[3], const float rd[3], const float inv[3], float t_limit) {
    if (c.vs == 1.0f)
        return march_chunk_vs<skip_water, skip_glass, sun_ray, true>(p, f, s, c, ci, ro_w, rd, inv, t_limit);
    return march_chunk_vs<skip_water, skip_glass, sun_ray, false>(p, f, s, c, ci, ro_w, rd, inv, t_limit);
}

// depth_key / make_key (common.wgsl:1320-1330)
__device__ __forceinline__ uint32_t depth_key(float far_, float t) {
    float x = 1.0f - t / far_;
    x = fminf(fmaxf(x, 0.0f), 1.0f);
    return (uint32_t)(x * 16777215.0f);
}

__device__ __forceinline__ uint64_t make_key(float far_, const Hit &h, uint32_t ci) {
    uint32_t d = depth_key(far_, h.t);
    uint32_t vx = h.voxel & 63u, vy = (h.voxel >> 6) & 63u, vz = (h.voxel >> 12) & 63u;
    uint32_t mx = h.micro & 3u, my = (h.micro >> 2) & 3u, mz = (h.micro >> 4) & 3u;
    uint32_t b = ((vx | (mx << 6)) << 16) | ((vy | (my << 6)) << 8) | (vz | (mz << 6));
    return ((uint64_t)d << 40) | ((uint64_t)h.normal << 37) | ((uint64_t)ci << 24) | (uint64_t)b;
}

// shadow_ray_t (common.wgsl:1156-1195): chunk-grid DDA over at most 24 cells along the
// sun direction; its `inv = 1.0 / rd` is the frame's inv_sun.
__device__ inline float shadow_ray_t(const Pools &p, const FrameK &f, const Spec &s, const float ro[3],
                                     float max_dist) {
    const float *rd = f.sun, *inv = f.inv_sun;
    bool pd[3];
    int32_t cell[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        pd[i] = rd[i] > 0.0f;
        cell[i] = (int32_t)floorf(ro[i] / 64.0f);
    }
    uint32_t last = NO_CHUNK;
    for (uint32_t i = 0; i < 24u; i++) {
        bool ok;
        uint32_t gi = grid_index(f, cell[0], cell[1], cell[2], ok);
        if (!ok) return -1.0f;
        uint32_t ci = ld_grid(p, f, gi);
        if (ci != NO_CHUNK && ci != last) {
            last = ci;
            if (ci < p.n_chunks) {
                ChunkK c = ld_chunk(p, ci);
                Hit m = march_chunk<true, true, true>(p, f, s, c, ci, ro, rd, inv, max_dist);
                if (m.hit) return m.t;
            }
        }
        float tn[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float bound = (float)(pd[k] ? cell[k] + 1 : cell[k]) * 64.0f;
            tn[k] = (bound - ro[k]) * inv[k];
        }
        uint32_t axis = 0;
        float t;
        if (tn[0] <= tn[1] && tn[0] <= tn[2]) t = tn[0];
        else if (tn[1] <= tn[2]) { t = tn[1]; axis = 1; }
        else { t = tn[2]; axis = 2; }
        if (t >= max_dist) return -1.0f;
        int32_t step = pd[axis] ? 1 : -1;
        if (axis == 0) cell[0] += step;
        else if (axis == 1) cell[1] += step;
        else cell[2] += step;
    }
    return -1.0f;
}

// shadow_ray (common.wgsl:1098-1148): a rising ray that starts under the sea level jumps to
// the sea surface first (SPEC_WATER_REPAIR), then walks the chunk grid as shadow_ray_t
// does from there; any hit blocks. shadow_ray_t's hit distance is march_chunk's
// max(t_enter, 0) * voxel_size, never negative, so ">= 0" is the walk's hit flag.
__device__ inline bool shadow_ray(const Pools &p, const FrameK &f, const Spec &s, const float ro[3], float max_dist) {
    const float *rd = f.sun;
    float r_ro[3] = {ro[0], ro[1], ro[2]};
    float r_dist = max_dist;
    if ((s.bits & SPEC_WATER_REPAIR) && rd[1] > 1e-4f && ro[1] < f.sea_level) {
        const float t_surf = (f.sea_level - ro[1]) / rd[1];
        if (t_surf >= r_dist) return false;
#pragma unroll
        for (int k = 0; k < 3; k++) r_ro[k] = ro[k] + rd[k] * t_surf;
        r_dist = r_dist - t_surf;
    }
    return shadow_ray_t(p, f, s, r_ro, r_dist) >= 0.0f;
}

// hit_t (common.wgsl:1300-1318)
__device__ __forceinline__ float hit_t(const FrameK &f, const ChunkK &c, uint32_t vx, uint32_t vy, uint32_t vz,
                                       uint32_t mx, uint32_t my, uint32_t mz, uint32_t nid, const float rd[3]) {
    float bx = c.origin[0] + (float)vx * c.vs, by = c.origin[1] + (float)vy * c.vs, bz = c.origin[2] + (float)vz * c.vs;
    if (nid >= NORMAL_CROSS_A) {
        if (nid == NORMAL_CROSS_A) return ((bx - bz) - (f.cam[0] - f.cam[2])) / (rd[0] - rd[2]);
        return ((bx + bz + c.vs) - (f.cam[0] + f.cam[2])) / (rd[0] + rd[2]);
    }
    uint32_t axis = nid >> 1;
    float ms = c.vs / 4.0f;
    uint32_t add = (nid & 1u) ? 1u : 0u;
    float plane;
    if (axis == 0) plane = bx + (float)(mx + add) * ms;
    else if (axis == 1) plane = by + (float)(my + add) * ms;
    else plane = bz + (float)(mz + add) * ms;
    const float cam_a = axis == 0 ? f.cam[0] : (axis == 1 ? f.cam[1] : f.cam[2]);
    const float rd_a = axis == 0 ? rd[0] : (axis == 1 ? rd[1] : rd[2]);
    return (plane - cam_a) / rd_a;
}

// f32 -> f16 bits, IEEE round-to-nearest-even (the R16F store of primary_shadow).
__device__ __forceinline__ uint16_t f32_to_f16_bits(float x) {
    _Float16 h = (_Float16)x;
    return __builtin_bit_cast(uint16_t, h);
}

// normal_of (common.wgsl:387-398)
__device__ __forceinline__ void normal_of(uint32_t id, float n[3]) {
    if (id >= NORMAL_CROSS_A) {
        float s = (id == NORMAL_CROSS_B) ? 1.0f : -1.0f;
        n[0] = 0.70710678f; n[1] = 0.0f; n[2] = s * 0.70710678f;
        return;
    }
    uint32_t axis = id >> 1;
    float s = (id & 1u) ? 1.0f : -1.0f;
    n[0] = axis == 0 ? s : 0.0f;
    n[1] = axis == 1 ? s : 0.0f;
    n[2] = axis == 2 ? s : 0.0f;
}

// shadow_guide (shadow.wgsl:14-30): whether the pixel's key gets a sun ray, and its
// origin hit + n * 0.02 * vs, with hit = cam + rd * hit_t.
// rd is the pixel's ray_dir(px + 0.5, py + 0.5), which the caller has already computed.
__device__ inline bool shadow_origin(const Pools &p, const FrameK &f, const Spec &s, uint64_t key, const float rd[3],
                                     float o[3]) {
    const bool active = (f.flags & VX_FLAG_SHADOWS) && !(f.flags & VX_FLAG_HEATMAP) && f.daylight > 0.01f;
    if (!active || key == 0ull) return false;
    uint32_t ci = (uint32_t)(key >> 24) & 0x1FFFu;
    uint32_t low = (uint32_t)(key & 0xFFFFFFull);
    uint32_t vx = (low >> 16) & 63u, vy = (low >> 8) & 63u, vz = low & 63u;
    if (ci >= p.n_chunks) return false;
    ChunkK c = ld_chunk(p, ci);
    uint32_t id = block_at(p, c, vx, vy, vz);
    if (id == WATER_ID || ((s.bits & SPEC_GLASS) && id == GLASS_ID)) return false;
    uint32_t nid = (uint32_t)(key >> 37) & 7u;
    uint32_t mx = (low >> 22) & 3u, my = (low >> 14) & 3u, mz = (low >> 6) & 3u;
    float n[3];
    normal_of(nid, n);
    if (nid >= NORMAL_CROSS_A && (n[0] * rd[0] + n[1] * rd[1] + n[2] * rd[2]) > 0.0f) {
        n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2];
    }
    if (!((n[0] * f.sun[0] + n[1] * f.sun[1] + n[2] * f.sun[2]) > 0.0f)) return false;
    float t = hit_t(f, c, vx, vy, vz, mx, my, mz, nid, rd);
    if (!(t >= 0.0f)) return false;
    float e = 0.02f * c.vs;
#pragma unroll
    for (int i = 0; i < 3; i++) o[i] = (f.cam[i] + rd[i] * t) + n[i] * e;
    return true;
}

// primary_shadow's stored value (shadow.wgsl:34-48): R16F of clamp(t, 0.001, 65504) for
// a blocked sun ray, 0 otherwise.
__device__ __forceinline__ uint16_t shadow_word(float blocker) {
    float distance = 0.0f;
    if (blocker >= 0.0f) distance = fminf(fmaxf(blocker, 0.001f), 65504.0f);
    return f32_to_f16_bits(distance);
}

// primary_shadow for one pixel (shadow.wgsl:8-48). Returns the R16F bits and, through
// traced, whether a sun ray was cast (the pixel's shadow_guide was valid).
__device__ inline uint16_t shadow_pixel(const Pools &p, const FrameK &f, const Spec &s, uint64_t key, const float rd[3],
                                        bool &traced) {
    float o[3];
    traced = shadow_origin(p, f, s, key, rd, o);
    return shadow_word(traced ? shadow_ray_t(p, f, s, o, f.shadow_dist) : -1.0f);
}

// box_outside / box_distance (tile_select.wgsl:1-9)
__device__ __forceinline__ bool box_outside(const float bmin[3], const float bmax[3], const float n[3],
                                            const float cam[3]) {
    float q0 = (n[0] > 0.0f ? bmax[0] : bmin[0]) - cam[0];
    float q1 = (n[1] > 0.0f ? bmax[1] : bmin[1]) - cam[1];
    float q2 = (n[2] > 0.0f ? bmax[2] : bmin[2]) - cam[2];
    return (n[0] * q0 + n[1] * q1 + n[2] * q2) < 0.0f;
}

__device__ __forceinline__ float box_distance(const float p[3], const float bmin[3], const float bmax[3]) {
    float q[3];
#pragma unroll
    for (int i = 0; i < 3; i++) q[i] = fmaxf(fmaxf(bmin[i] - p[i], p[i] - bmax[i]), 0.0f);
    return sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
}

// The per-tile culling set-up of tile_select (tile_select.wgsl:16-57).
struct TileCull {
    float n[4][3];
    float reach;
};

__device__ inline TileCull tile_cull_setup(const FrameK &f, uint32_t gx, uint32_t gy) {
    TileCull tc;
    float x0 = (float)(gx * 8u), y0 = (float)(gy * 8u);
    float x1 = x0 + 8.0f, y1 = y0 + 8.0f;
    float d[4][3], ce[3];
    ray_dir(f, x0, y0, d[0][0], d[0][1], d[0][2]);
    ray_dir(f, x1, y0, d[1][0], d[1][1], d[1][2]);
    ray_dir(f, x1, y1, d[2][0], d[2][1], d[2][2]);
    ray_dir(f, x0, y1, d[3][0], d[3][1], d[3][2]);
    ray_dir(f, (x0 + x1) * 0.5f, (y0 + y1) * 0.5f, ce[0], ce[1], ce[2]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float *a = d[k];
        const float *b = d[(k + 1) & 3];
        float cx = a[1] * b[2] - a[2] * b[1];
        float cy = a[2] * b[0] - a[0] * b[2];
        float cz = a[0] * b[1] - a[1] * b[0];
        bool keep = (cx * ce[0] + cy * ce[1] + cz * ce[2]) >= 0.0f;
        tc.n[k][0] = keep ? cx : -cx;
        tc.n[k][1] = keep ? cy : -cy;
        tc.n[k][2] = keep ? cz : -cz;
    }
    tc.reach = f.far_;
    const uint32_t uw = VX_FLAG_UNDERWATER | VX_FLAG_WATER_FAR;
    if ((f.flags & uw) == uw) {
        float up = fmaxf(fmaxf(d[0][1], d[1][1]), fmaxf(d[2][1], d[3][1]));
        float submerged = fmaxf(f.sea_level - f.cam[1], 0.0f);
        float escape = 1e30f;
        if (up > 0.0f) escape = submerged / up;
        if (escape >= 128.0f) {
            tc.reach = fminf(f.far_, 128.0f);
            float dep_far = submerged - up * 128.0f;
            if ((f.flags & VX_FLAG_WATER_DARK) && fminf(submerged, dep_far) >= 15.0f) tc.reach = fminf(tc.reach, 64.0f);
        }
    }
    return tc;
}

// tile_cull_setup with the five ray_dir evaluations spread over lanes 0-4 of the wave
// (corner k on lane k, the centre on lane 4) and broadcast: the same function on the same
// arguments, so the same planes, for one ray_dir's worth of instructions instead of five.
// Needs the whole wave active.
__device__ inline TileCull tile_cull_setup_wave(const FrameK &f, uint32_t gx, uint32_t gy, uint32_t lane) {
    TileCull tc;
    const float x0 = (float)(gx * 8u), y0 = (float)(gy * 8u);
    const float x1 = x0 + 8.0f, y1 = y0 + 8.0f;
    // corners in the order (x0,y0) (x1,y0) (x1,y1) (x0,y1), then the centre
    const float qx = lane == 4u ? (x0 + x1) * 0.5f : ((lane == 1u || lane == 2u) ? x1 : x0);
    const float qy = lane == 4u ? (y0 + y1) * 0.5f : ((lane == 2u || lane == 3u) ? y1 : y0);
    float q[3];
    ray_dir(f, qx, qy, q[0], q[1], q[2]);
    float d[4][3], ce[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int k = 0; k < 4; k++) d[k][i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q[i]), k));
        ce[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q[i]), 4));
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float *a = d[k];
        const float *b = d[(k + 1) & 3];
        float cx = a[1] * b[2] - a[2] * b[1];
        float cy = a[2] * b[0] - a[0] * b[2];
        float cz = a[0] * b[1] - a[1] * b[0];
        bool keep = (cx * ce[0] + cy * ce[1] + cz * ce[2]) >= 0.0f;
        tc.n[k][0] = keep ? cx : -cx;
        tc.n[k][1] = keep ? cy : -cy;
        tc.n[k][2] = keep ? cz : -cz;
    }
    tc.reach = f.far_;
    const uint32_t uw = VX_FLAG_UNDERWATER | VX_FLAG_WATER_FAR;
    if ((f.flags & uw) == uw) {
        float up = fmaxf(fmaxf(d[0][1], d[1][1]), fmaxf(d[2][1], d[3][1]));
        float submerged = fmaxf(f.sea_level - f.cam[1], 0.0f);
        float escape = 1e30f;
        if (up > 0.0f) escape = submerged / up;
        if (escape >= 128.0f) {
            tc.reach = fminf(f.far_, 128.0f);
            float dep_far = submerged - up * 128.0f;
            if ((f.flags & VX_FLAG_WATER_DARK) && fminf(submerged, dep_far) >= 15.0f) tc.reach = fminf(tc.reach, 64.0f);
        }
    }
    return tc;
}

// Classification of one (tile, chunk) pair by tile_select (tile_select.wgsl:59-91):
// 0 culled, 1 selected, 2 deferred by Hi-Z. dmin receives box_distance when computed.
__device__ __forceinline__ uint32_t classify(const FrameK &f, const TileCull &tc, const float bmin[3],
                                             const float bmax[3], uint32_t ci, uint32_t tile_far, bool use_hiz,
                                             float &dmin) {
    dmin = 0.0f;
    if (ci == f.camera_chunk) return 1u;
    if (box_outside(bmin, bmax, f.fwd, f.cam)) return 0u;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (box_outside(bmin, bmax, tc.n[k], f.cam)) return 0u;
    dmin = box_distance(f.cam, bmin, bmax);
    if (dmin >= tc.reach) return 0u;
    if (use_hiz && depth_key(f.far_, dmin) < tile_far) return 2u;
    return 1u;
}

// ---- Secondary legs (resolve.wgsl shade_water / shade_glass, flat water normals) -------
constexpr float WATER_IOR = 1.333f, SURFACE_EPS = 0.02f;               // common.wgsl:1256-1258
constexpr float WATER_REFRACT_NEAR = 96.0f, WATER_REFRACT_FAR = 192.0f;  // resolve.wgsl:873-886
constexpr float WATER_REFLECT_NEAR = 512.0f, WATER_REFLECT_FAR = 1024.0f;
constexpr float WATER_REFRACT_DIST = 96.0f, WATER_REFLECT_DIST = 768.0f, GLASS_TRANSMIT_DIST = 192.0f;
constexpr uint32_t FLAG_WATER_REFLECT = 16u, FLAG_WATER_REFRACT = 32u;  // common.wgsl:146-147
constexpr uint32_t FLAG_WATER_SHADOW_CUT = 262144u;                     // src/render/mod.rs:198
// Sun-ray range of shade_hit at a leg's hit: SPEC_WATER_SHADOW_DIST (src/render/mod.rs:
// 325-327, :2427-2432) for the refract leg, frame.shadow_dist for the water reflect leg,
// GLASS_SHADOW_DIST (resolve.wgsl:888) for the glass legs.
constexpr float WATER_SHADOW_DIST = 16.0f, WATER_SHADOW_DIST_PRE41 = 48.0f, GLASS_SHADOW_DIST = 48.0f;
constexpr uint32_t SEC_SUN_TESTED = 1u << 29, SEC_SUN_BLOCKED = 1u << 28;

struct WorldHit {
    bool hit;
    float t;
    uint32_t normal, voxel, ci;
};

__device__ __forceinline__ float wgsl_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// trace_world (common.wgsl:1205-1253): chunk-grid DDA over at most 16 cells from ro,
// marching each resident chunk with skip_glass; the first hit ends the walk. The chunk
// march takes 1 / rd of the leg itself (march_chunk's own `let inv = 1.0 / rd`); the grid
// walk steps with the tiny-component guard of trace_world.
template <bool kSkipWater>
__device__ inline WorldHit trace_world(const Pools &p, const FrameK &f, const Spec &s, const float ro[3],
                                       const float rd[3], float max_dist) {
    WorldHit out{false, max_dist, 0u, 0u, NO_CHUNK};
    float inv[3], inv_rd[3];
    bool pos_dir[3];
    int32_t cell[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const bool tiny = fabsf(rd[k]) < 1e-6f;
        const float dir = tiny ? 1e-6f * wgsl_sign(rd[k] + 1e-9f) : rd[k];
        inv[k] = 1.0f / dir;
        inv_rd[k] = 1.0f / rd[k];
        pos_dir[k] = dir > 0.0f;
        cell[k] = (int32_t)floorf(ro[k] / 64.0f);
    }
    for (uint32_t i = 0; i < 16u; i++) {
        bool ok;
        const uint32_t gi = grid_index(f, cell[0], cell[1], cell[2], ok);
        if (!ok) return out;
        const uint32_t ci = ld_grid(p, f, gi);
        float tn[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float bound = (float)(pos_dir[k] ? cell[k] + 1 : cell[k]) * 64.0f;
            tn[k] = (bound - ro[k]) * inv[k];
        }
        float t_exit = tn[0];
        uint32_t axis = 0;
        if (tn[1] < t_exit) { t_exit = tn[1]; axis = 1; }
        if (tn[2] < t_exit) { t_exit = tn[2]; axis = 2; }
        if (ci != NO_CHUNK && ci < p.n_chunks) {
            const ChunkK c = ld_chunk(p, ci);
            const Hit h = march_chunk<kSkipWater, true, false>(p, f, s, c, ci, ro, rd, inv_rd, fminf(t_exit, max_dist));
            if (h.hit) {
                out.hit = true;
                out.t = h.t;
                out.normal = h.normal;
                out.voxel = h.voxel;
                out.ci = ci;
                return out;
            }
        }
        if (t_exit >= max_dist) return out;
        const int32_t step = pos_dir[axis] ? 1 : -1;
        if (axis == 0) cell[0] += step;
        else if (axis == 1) cell[1] += step;
        else cell[2] += step;
    }
    return out;
}

// WGSL builtins with a fixed operation order: dot = (x*x' + y*y') + z*z'.
__device__ __forceinline__ float dot3o(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
// refract(I, N, eta): k = 1 - eta*eta*(1 - dot(N,I)^2); k < 0 ? 0 : eta*I - (eta*dot(N,I) + sqrt(k))*N
__device__ __forceinline__ void wgsl_refract(const float i[3], const float n[3], float eta, float out[3]) {
    const float d = dot3o(n, i);
    const float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) {
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    const float sc = eta * d + sqrtf(k);
#pragma unroll
    for (int q = 0; q < 3; q++) out[q] = eta * i[q] - sc * n[q];
}
// reflect(I, N) = I - 2*dot(N,I)*N
__device__ __forceinline__ void wgsl_reflect(const float i[3], const float n[3], float out[3]) {
    const float d2 = 2.0f * dot3o(n, i);
#pragma unroll
    for (int q = 0; q < 3; q++) out[q] = i[q] - d2 * n[q];
}
// smoothstep(e0, e1, x) = t*t*(3 - 2t), t = clamp((x - e0) / (e1 - e0), 0, 1)
// The Snell bend of march (march.wgsl:23-45): an underwater camera's ray that misses the
// pair's chunk going up is refracted at the sea surface (from water, eta = WATER_IOR,
// about (0, -1, 0)) and marched again through the same chunk from just past the surface,
// skipping water, within far - t_surf; a hit there replaces the miss with t = t_surf +
// its distance, and the iterations add up.
__device__ inline void snell_bend(const Pools &p, const FrameK &f, const Spec &s, const ChunkK &c, uint32_t ci,
                                  const float rd[3], Hit &h) {
    const float depth = f.sea_level - f.cam[1];
    if (!(depth > 0.0f)) return;
    const float t_surf = depth / rd[1];
    if (!(t_surf < f.far_)) return;
    const float n[3] = {0.0f, -1.0f, 0.0f};
    float rd2[3], ro2[3], inv2[3];
    wgsl_refract(rd, n, WATER_IOR, rd2);
    if (!(dot3o(rd2, rd2) > 0.0f)) return;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        ro2[k] = (f.cam[k] + rd[k] * t_surf) + rd2[k] * SURFACE_EPS;
        inv2[k] = 1.0f / rd2[k];
    }
    const Hit h2 = march_chunk<true, false, false>(p, f, s, c, ci, ro2, rd2, inv2, f.far_ - t_surf);
    if (h2.hit) {
        const uint32_t straight = h.iters;
        h = h2;
        h.t = t_surf + h2.t;
        h.iters = straight + h2.iters;
    }
}

__device__ __forceinline__ float wgsl_smoothstep(float e0, float e1, float x) {
    const float t = fminf(fmaxf((x - e0) / (e1 - e0), 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// ---- waved water (water_ctx, resolve.wgsl:1078-1167) ---------------------------------------
constexpr uint32_t NORMAL_PY = 3u;  // common.wgsl:295
constexpr uint32_t FLAG_WAVES = 512u, FLAG_WAVE_ANISO = 1024u, FLAG_WAVE_SHOAL = 2048u,
                   FLAG_WAVE_FILL = 4096u;  // src/render/mod.rs:180-186
constexpr float WAVE_TAU = 6.28318531f, WAVE_GRAZE_MIN = 1e-3f, SHOAL_DEPTH = 6.0f, SHOAL_MAX = 6.0f,
                WAVE_FILTER_K = 1.0f;  // resolve.wgsl:902-912

// cos and exp of the wave field. WGSL leaves their accuracy to the implementation; these
// are fixed f32 steps, the ones the oracle takes:
//   cos: k = round-to-even(x * 2/pi); r = ((x - k*A) - k*B) - k*C with A + B + C = pi/2
//        (Cody and Waite's reduction: A and B have 12 significant bits, so k*A and k*B
//        are exact for |k| < 4096); the Taylor polynomial of cos (degree 8) or sin
//        (degree 9) on |r| <= pi/4, chosen and signed by k mod 4, by Horner's rule;
//   exp: 0 for x < -86; k = round-to-even(x / ln 2), r = (x - k*L1) - k*L2 (L1 of 12
//        bits); the degree-7 Taylor polynomial of exp(r), times 2^k.
__device__ __forceinline__ float wave_cos(float x) {
    const float kf = rintf(x * 0x1.45f306p-1f);
    const float r = ((x - kf * 0x1.922p0f) - kf * -0x1.2aep-18f) - kf * -0x1.de973ep-31f;
    const float r2 = r * r;
    float cs = 0x1.a01a02p-16f;
    cs = cs * r2 - 0x1.6c16c2p-10f;
    cs = cs * r2 + 0x1.555556p-5f;
    cs = cs * r2 - 0.5f;
    cs = cs * r2 + 1.0f;
    float sn = 0x1.71de3ap-19f;
    sn = sn * r2 - 0x1.a01a02p-13f;
    sn = sn * r2 + 0x1.111112p-7f;
    sn = sn * r2 - 0x1.555556p-3f;
    sn = sn * r2 + 1.0f;
    sn = sn * r;
    const int32_t q = (int32_t)kf & 3;
    return q == 0 ? cs : (q == 1 ? -sn : (q == 2 ? -cs : sn));
}

__device__ __forceinline__ float wave_exp(float x) {
    if (x < -86.0f) return 0.0f;
    const float kf = rintf(x * 0x1.715476p0f);
    const float r = (x - kf * 0x1.62ep-1f) - kf * 0x1.0bfbe8p-15f;
    float e = 0x1.a01a02p-13f;
    e = e * r + 0x1.6c16c2p-10f;
    e = e * r + 0x1.111112p-7f;
    e = e * r + 0x1.555556p-5f;
    e = e * r + 0x1.555556p-3f;
    e = e * r + 0.5f;
    e = e * r + 1.0f;
    e = e * r + 1.0f;
    return ldexpf(e, (int32_t)kf);
}

// normalize(v) = v / sqrt(dot(v, v)) (common conventions of the restatement)
__device__ __forceinline__ void wgsl_normalize(float v[3]) {
    const float len = sqrtf(dot3o(v, v));
#pragma unroll
    for (int k = 0; k < 3; k++) v[k] = v[k] / len;
}

// wave_octave (resolve.wgsl:928-944)
__device__ __forceinline__ void wave_octave(const float p[2], float dx, float dz, float lambda, float slope,
                                            float phase, float px, const float vh[2], float kk, bool aniso,
                                            float out[3]) {
    float fade;
    if (aniso) {
        const float d = dx * vh[0] + dz * vh[1];
        const float ext = px * sqrtf(1.0f + kk * d * d);
        const float u = ext / lambda;
        fade = wave_exp(-WAVE_FILTER_K * u * u);
    } else {
        fade = wgsl_smoothstep(2.0f, 4.0f, lambda / fmaxf(px, 1e-4f));
    }
    const float g = slope * fade * wave_cos(WAVE_TAU / lambda * (p[0] * dx + p[1] * dz) + phase);
    out[0] = dx * g;
    out[1] = dz * g;
    out[2] = 0.5f * slope * slope * (1.0f - fade * fade);
}

// wave_field (resolve.wgsl:946-967): the octaves summed left to right; the first octave's
// wavelength is wave_scale itself.
__device__ inline void wave_field(const FrameK &f, const Spec &s, float hx, float hz, float px_world,
                                  const float vh[2], float kk, bool aniso, float shoal, float out[3]) {
    constexpr float DD[4][2] = {{0.913089f, 0.407760f}, {-0.388685f, 0.921371f}, {-0.980382f, -0.197108f},
                                {0.477328f, -0.878725f}};
    constexpr float DDIV[4] = {0.0f, 2.17f, 4.71f, 10.20f}, DAMP[4] = {0.42f, 0.27f, 0.19f, 0.12f},
                    DPH[4] = {1.0000f, 1.4731f, 2.1703f, 3.1937f};
    constexpr float FD[8][2] = {{-0.438371f, 0.898794f}, {-0.927184f, 0.374607f}, {-0.190809f, 0.981627f},
                                {-0.669131f, 0.743145f}, {0.121869f, 0.992546f},  {0.694658f, 0.719340f},
                                {0.970296f, 0.241922f},  {0.438371f, 0.898794f}};
    constexpr float FDIV[8] = {0.0f, 1.3934f, 1.9417f, 2.7056f, 3.7700f, 5.2533f, 7.3201f, 10.2000f},
                    FAMP[8] = {0.3086f, 0.2584f, 0.2163f, 0.1810f, 0.1515f, 0.1269f, 0.1062f, 0.0889f},
                    FPH[8] = {1.0000f, 1.1804f, 1.3934f, 1.6449f, 1.9417f, 2.2920f, 2.7056f, 3.1937f};
    const float p[2] = {hx, hz};
    const float l0 = f.wave_scale;
    const float a = f.wave_amp * shoal;
    const float ph = f.time * f.wave_speed * WAVE_TAU;
    const bool fill = (s.bits & SPEC_WAVE_FILL) && (f.flags & FLAG_WAVE_FILL);
    float o[3];
    if (fill) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float lambda = i == 0 ? l0 : l0 / FDIV[i];
            wave_octave(p, FD[i][0], FD[i][1], lambda, a * FAMP[i], ph * FPH[i], px_world, vh, kk, aniso, o);
            if (i == 0) { out[0] = o[0]; out[1] = o[1]; out[2] = o[2]; }
            else { out[0] = out[0] + o[0]; out[1] = out[1] + o[1]; out[2] = out[2] + o[2]; }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float lambda = i == 0 ? l0 : l0 / DDIV[i];
            wave_octave(p, DD[i][0], DD[i][1], lambda, a * DAMP[i], ph * DPH[i], px_world, vh, kk, aniso, o);
            if (i == 0) { out[0] = o[0]; out[1] = o[1]; out[2] = o[2]; }
            else { out[0] = out[0] + o[0]; out[1] = out[1] + o[1]; out[2] = out[2] + o[2]; }
        }
    }
}

// water_ctx's normals (resolve.wgsl:1078-1167): n the face normal; on waved top faces nw
// the wave normal and nt the wave normal at the reflect slope (the shoal depth from a
// short trace straight down scales the waves) and gloss from the unresolved variance;
// nw and nt blended back toward n by the optical roughness of that variance. Water-look
// ripples are refused.
struct WaterNormals {
    float n[3], nw[3], nt[3], gloss, rough;
    bool waved;
};
__device__ inline void water_normals(const Pools &p, const FrameK &f, const Spec &s, uint32_t nid, const float hit[3],
                                     const float rd[3], float t, WaterNormals &o) {
    normal_of(nid, o.n);
#pragma unroll
    for (int k = 0; k < 3; k++) o.nw[k] = o.nt[k] = o.n[k];
    o.gloss = 1.0f;
    float unresolved = 0.0f;
    o.waved = (s.bits & SPEC_WAVES) && (f.flags & FLAG_WAVES) && nid == NORMAL_PY;
    if (o.waved) {
        const float px_world = fmaxf(t, 1e-3f) * 2.0f * f.tan_half_fov / (float)f.res[1];
        float kk = 0.0f, vh[2] = {0.0f, 0.0f};
        const bool aniso = (s.bits & SPEC_WAVE_ANISO) && (f.flags & FLAG_WAVE_ANISO);
        if (aniso) {
            const float k = 1.0f / fmaxf(fabsf(rd[1]), WAVE_GRAZE_MIN);
            kk = k * k - 1.0f;
            const float is = 1.0f / sqrtf(fmaxf(rd[0] * rd[0] + rd[2] * rd[2], 1e-12f));
            vh[0] = rd[0] * is;
            vh[1] = rd[2] * is;
        }
        float shoal = 1.0f;
        if ((s.bits & SPEC_WAVE_SHOAL) && (f.flags & FLAG_WAVE_SHOAL)) {
            const float dro[3] = {floorf(hit[0]) + 0.5f, hit[1] - SURFACE_EPS, floorf(hit[2]) + 0.5f};
            const float sdir[3] = {0.001f, -0.999999f, 0.001f};
            const WorldHit dh = trace_world<true>(p, f, s, dro, sdir, SHOAL_MAX);
            shoal = wgsl_smoothstep(0.0f, SHOAL_DEPTH, dh.hit ? dh.t : SHOAL_MAX);
        }
        float w[3];
        wave_field(f, s, hit[0], hit[2], px_world, vh, kk, aniso, shoal, w);
        unresolved = fmaxf(w[2], 0.0f);
        o.nw[0] = -w[0];
        o.nw[1] = 1.0f;
        o.nw[2] = -w[1];
        wgsl_normalize(o.nw);
        const float len = sqrtf(w[0] * w[0] + w[1] * w[1]);
        const float m = fminf(1.0f, f.wave_reflect_slope / fmaxf(len, 1e-6f));
        o.nt[0] = -(w[0] * m);
        o.nt[1] = 1.0f;
        o.nt[2] = -(w[1] * m);
        wgsl_normalize(o.nt);
        o.gloss = 1.0f / (1.0f + 160.0f * w[2]);  // 0.5 * GLINT_POWER
    }
    const float rough = fminf(fmaxf(1.0f - wave_exp(-8.0f * unresolved), 0.0f), 1.0f);
    o.rough = rough;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        o.nw[k] = o.nw[k] * (1.0f - rough) + o.n[k] * rough;
        o.nt[k] = o.nt[k] * (1.0f - rough) + o.n[k] * rough;
    }
    wgsl_normalize(o.nw);
    wgsl_normalize(o.nt);
}

// water_ctx's reflect-leg direction mirror_t: reflect(rd, nt), on waved faces kept above
// the horizon.
__device__ inline void water_mirror_t(const Pools &p, const FrameK &f, const Spec &s, uint32_t nid, const float hit[3],
                                      const float rd[3], float t, float out[3]) {
    WaterNormals w;
    water_normals(p, f, s, nid, hit, rd, t, w);
    wgsl_reflect(rd, w.nt, out);
    if (w.waved) {
        out[1] = fmaxf(out[1], 0.0f);
        wgsl_normalize(out);
    }
}

// One secondary leg to trace: origin, direction, range, and which leg of the pixel.
struct SecLeg {
    float ro[3], rd[3], max_dist;
    float sun_dist;  // range of the sun ray at the leg's hit
    bool water;      // water legs skip water voxels (trace_world(..., true))
};

// The sun test shade_hit makes at a leg's hit (shade_hit_compact, resolve.wgsl:587-595 for
// n and vs, :741-745; shade_hit_legacy :507-514): n = normal_of(normal), turned toward the
// ray for cross quads; only when dot(n, sun) > 0, daylight > 0.01 and FLAG_SHADOWS does
// shadow_ray run, from hit + n * (0.02 * vs). A glass leg that ends on water is shaded by
// shade_water (resolve.wgsl:1284-1289, :1305-1309): no test at that hit (the transmit
// leg's water legs are records 2 and 3, tested at their own hits).
__device__ inline uint32_t leg_sun(const Pools &p, const FrameK &f, const Spec &s, const SecLeg &l,
                                   const WorldHit &w, const ChunkK &c, uint32_t material) {
    if (!w.hit || (!l.water && material == WATER_ID)) return 0u;
    float n[3], o[3];
    normal_of(w.normal, n);
    if (w.normal >= NORMAL_CROSS_A && dot3o(n, l.rd) > 0.0f) {
#pragma unroll
        for (int k = 0; k < 3; k++) n[k] = -n[k];
    }
    const float ndl = dot3o(n, f.sun);
    if (!(ndl > 0.0f && f.daylight > 0.01f) || !(f.flags & VX_FLAG_SHADOWS)) return 0u;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float hit = l.ro[k] + l.rd[k] * w.t;
        o[k] = hit + n[k] * (0.02f * c.vs);
    }
    return SEC_SUN_TESTED | (shadow_ray(p, f, s, o, l.sun_dist) ? SEC_SUN_BLOCKED : 0u);
}

// shade_water's two legs from a water hit (resolve.wgsl:1209-1220): the refract leg about
// the face normal (water_refract_leg, :1029-1042) in legs[0], the reflect leg along
// water_ctx's mirror_t (water_reflect_leg, :1044-1058) in legs[1]. kDirs false: only which
// legs exist (want) and their origins' inputs are needed, not the reflect direction.
template <bool kDirs>
__device__ inline void water_leg_pair(const Pools &p, const FrameK &f, const Spec &s, uint32_t nid, const float hit[3],
                                      const float rd[3], float t, SecLeg legs[2], bool want[2]) {
    float n[3];
    normal_of(nid, n);
    const float refract_detail = 1.0f - wgsl_smoothstep(WATER_REFRACT_NEAR, WATER_REFRACT_FAR, t);
    const float reflect_detail = 1.0f - wgsl_smoothstep(WATER_REFLECT_NEAR, WATER_REFLECT_FAR, t);
    if ((f.flags & FLAG_WATER_REFRACT) && refract_detail > 0.0f) {
        SecLeg &l = legs[0];
        wgsl_refract(rd, n, 1.0f / WATER_IOR, l.rd);
#pragma unroll
        for (int k = 0; k < 3; k++) l.ro[k] = hit[k] + l.rd[k] * SURFACE_EPS;
        l.max_dist = WATER_REFRACT_DIST;
        l.sun_dist = (f.flags & FLAG_WATER_SHADOW_CUT) ? WATER_SHADOW_DIST : WATER_SHADOW_DIST_PRE41;
        l.water = true;
        want[0] = true;
    }
    if ((f.flags & FLAG_WATER_REFLECT) && reflect_detail > 0.0f) {
        SecLeg &l = legs[1];
        if constexpr (kDirs) water_mirror_t(p, f, s, nid, hit, rd, t, l.rd);
#pragma unroll
        for (int k = 0; k < 3; k++) l.ro[k] = hit[k] + n[k] * SURFACE_EPS;
        l.max_dist = WATER_REFLECT_DIST;
        l.sun_dist = f.shadow_dist;
        l.water = true;
        want[1] = true;
    }
}

// The primary hit a key names: ray direction, hit point, distance (hit_t), face and block.
struct KeyHit {
    float rd[3], hit[3], t;
    uint32_t nid, id;
    bool ok;
};
__device__ inline KeyHit key_hit(const Pools &p, const FrameK &f, uint64_t key, uint32_t px, uint32_t py) {
    KeyHit h;
    h.ok = false;
    if (key == 0ull) return h;
    const uint32_t ci = (uint32_t)(key >> 24) & 0x1FFFu;
    if (ci >= p.n_chunks) return h;
    const uint32_t low = (uint32_t)(key & 0xFFFFFFull);
    const uint32_t vx = (low >> 16) & 63u, vy = (low >> 8) & 63u, vz = low & 63u;
    const ChunkK c = ld_chunk(p, ci);
    h.id = block_at(p, c, vx, vy, vz);
    h.nid = (uint32_t)(key >> 37) & 7u;
    const uint32_t mx = (low >> 22) & 3u, my = (low >> 14) & 3u, mz = (low >> 6) & 3u;
    ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, h.rd[0], h.rd[1], h.rd[2]);
    h.t = hit_t(f, c, vx, vy, vz, mx, my, mz, h.nid, h.rd);
#pragma unroll
    for (int k = 0; k < 3; k++) h.hit[k] = f.cam[k] + h.rd[k] * h.t;
    h.ok = true;
    return h;
}

// The legs of one pixel's primary key (shade_water, resolve.wgsl:1078-1167, :1209-1220;
// shade_glass, :1261-1310): legs[0] = refract (water) / transmit (glass) when want[0],
// legs[1] = reflect. kDirs false: which legs exist, without the water reflect direction.
template <bool kDirs>
__device__ inline void secondary_legs(const Pools &p, const FrameK &f, const Spec &s, uint64_t key, uint32_t px,
                                      uint32_t py, SecLeg legs[2], bool want[2]) {
    want[0] = want[1] = false;
    const KeyHit h = key_hit(p, f, key, px, py);
    if (!h.ok) return;
    const bool glass = (s.bits & SPEC_GLASS) && h.id == GLASS_ID;
    if (h.id != WATER_ID && !glass) return;
    if (h.id == WATER_ID) {
        water_leg_pair<kDirs>(p, f, s, h.nid, h.hit, h.rd, h.t, legs, want);
        return;
    }
    float n[3];
    normal_of(h.nid, n);
    if (dot3o(n, h.rd) > 0.0f) {
#pragma unroll
        for (int k = 0; k < 3; k++) n[k] = -n[k];
    }
    const float dist = fmaxf(GLASS_TRANSMIT_DIST, f.far_);
    SecLeg &l = legs[0];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        l.ro[k] = h.hit[k] + h.rd[k] * SURFACE_EPS;
        l.rd[k] = h.rd[k];
    }
    l.max_dist = dist;
    l.sun_dist = GLASS_SHADOW_DIST;
    l.water = false;
    want[0] = true;
    if (s.bits & SPEC_GLASS_REFLECT) {
        SecLeg &m = legs[1];
#pragma unroll
        for (int k = 0; k < 3; k++) m.ro[k] = h.hit[k] + n[k] * SURFACE_EPS;
        wgsl_reflect(h.rd, n, m.rd);
        m.max_dist = dist;
        m.sun_dist = GLASS_SHADOW_DIST;
        m.water = false;
        want[1] = true;
    }
}

// The legs shade_water traces where a pixel's glass transmit leg ends on water
// (shade_glass, resolve.wgsl:1281-1289): from p2 = the leg's origin + rd * its distance,
// with the water face the leg hit and the distance t + the leg's; legs[0] refract, legs[1]
// reflect (the pixel's records 2 and 3). leg_t and leg_normal come from its record 0.
template <bool kDirs>
__device__ inline void glass_water_legs(const Pools &p, const FrameK &f, const Spec &s, uint64_t key, uint32_t px,
                                        uint32_t py, float leg_t, uint32_t leg_normal, SecLeg legs[2], bool want[2]) {
    want[0] = want[1] = false;
    const KeyHit h = key_hit(p, f, key, px, py);
    if (!h.ok) return;
    float p2[3];
#pragma unroll
    for (int k = 0; k < 3; k++) p2[k] = (h.hit[k] + h.rd[k] * SURFACE_EPS) + h.rd[k] * leg_t;
    water_leg_pair<kDirs>(p, f, s, leg_normal, p2, h.rd, h.t + leg_t, legs, want);
}

}  // namespace vxd
