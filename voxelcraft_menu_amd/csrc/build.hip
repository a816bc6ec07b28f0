// build.hip -- build_local (src/voxel/tree.rs:89-204) on the device, many chunks at once.
//
// Two passes over a batch of dense 64^3 chunks:
//   k_local_leaves    one 64-lane wave per L1 cell (16^3 voxels), one lane per leaf (4^3):
//                     the leaf's occupancy mask, its palette encoding (encode_leaf,
//                     src/voxel/attributes.rs:90-166) into a fixed 49-word slot, and the
//                     L1 summary (mask, full, dry, water line, flags, counts, bounds) by
//                     wave reductions;
//   k_local_assemble  one wave per chunk, one lane per L1 cell: prefix sums over the L1
//                     cells give leaf_start, leaf_prefix and the palette offsets, and the
//                     wave writes the chunk's flat vx_local_tree record (header, L1 nodes,
//                     kept leaf masks, attribute entries, palette words).
// The order of everything written follows build_local's loops (L1 cells, then leaves,
// in cell-bit order), so the record is the host build's, word for word.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ctx.hpp"

namespace vxd {

constexpr uint32_t VOL = 64u * 64u * 64u, LEAF_WORDS = 49u;
constexpr uint16_t B_AIR = 0, B_WATER = 13, B_GLASS = 10;
// foliage / cutout / emitter columns of the block table (src/block.rs:133-185), one bit
// per id; ids past the table read as its last row (reeds), as the host's lookup does.
constexpr uint32_t FOLIAGE_IDS = (1u << 18) | (1u << 24) | (1u << 25), CUTOUT_IDS = (1u << 7) | (1u << 16),
                   EMITTER_IDS = (1u << 11) | (1u << 20) | (1u << 21) | (1u << 22);
__device__ __forceinline__ bool id_flag(uint32_t mask, uint32_t id) { return (mask >> min(id, 25u)) & 1u; }

struct LeafInfo {
    uint64_t mask;
    uint32_t nwords;  // palette words of the leaf (0 for a single-id leaf)
    uint32_t entry;   // UNIFORM_BIT | id for a single-id leaf, else 0
    uint32_t lofs;    // palette offset of the leaf inside its L1 cell
    uint32_t pad;
};
struct L1Info {
    uint64_t mask;
    uint32_t full, dry, water_line, flags, emitters, pal_words, solid;
    uint32_t uni_state, uni_id;  // 0 empty, 1 one block id throughout (uni_id), 2 mixed
    uint32_t bmin, bmax;         // x | y << 8 | z << 16
};

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o);
    return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t lane) {
    uint32_t s = v;
#pragma unroll
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)s, d);
        if (lane >= d) s += o;
    }
    return s - v;
}

__global__ __launch_bounds__(64) void k_local_leaves(const uint16_t *__restrict__ dense, LeafInfo *__restrict__ leaves,
                                                     uint32_t *__restrict__ words, L1Info *__restrict__ l1s) {
    __shared__ uint16_t ids[64 * 64];  // [voxel bit][lane]
    __shared__ uint8_t idx[64 * 64];   // palette index, [voxel bit][lane]
    const uint32_t chunk = blockIdx.x >> 6, l1 = blockIdx.x & 63u, lane = threadIdx.x;
    const uint32_t l1x = l1 & 3u, l1y = (l1 >> 2) & 3u, l1z = l1 >> 4;
    const uint32_t lx = lane & 3u, ly = (lane >> 2) & 3u, lz = lane >> 4;
    const uint32_t bx = l1x * 16u + lx * 4u, by = l1y * 16u + ly * 4u, bz = l1z * 16u + lz * 4u;
    const uint16_t *src = dense + (size_t)chunk * VOL;
    uint64_t m = 0;
    bool water = false, dry = false, foliage = false, cutout = false, glass = false;
    uint32_t maxy = 0, emitters = 0;
    uint32_t mn[3] = {64u, 64u, 64u}, mx[3] = {0u, 0u, 0u};
    for (uint32_t vz = 0; vz < 4u; vz++)
        for (uint32_t vy = 0; vy < 4u; vy++) {
            const uint2 row = *(const uint2 *)(src + (bx | ((bz + vz) << 6) | ((by + vy) << 12)));
            const uint32_t r[4] = {row.x & 0xFFFFu, row.x >> 16, row.y & 0xFFFFu, row.y >> 16};
#pragma unroll
            for (uint32_t vx = 0; vx < 4u; vx++) {
                const uint32_t id = r[vx], vb = vx | (vy << 2) | (vz << 4);
                ids[vb * 64u + lane] = (uint16_t)id;
                if (id == B_AIR) continue;
                m |= 1ull << vb;
                water |= id == B_WATER;
                if (id != B_WATER) {
                    dry = true;
                    maxy = max(maxy, ly * 4u + vy);
                }
                foliage |= id_flag(FOLIAGE_IDS, id);
                cutout |= id_flag(CUTOUT_IDS, id);
                glass |= id == B_GLASS;
                emitters += id_flag(EMITTER_IDS, id) ? 1u : 0u;
                const uint32_t p[3] = {bx + vx, by + vy, bz + vz};
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    mn[k] = min(mn[k], p[k]);
                    mx[k] = max(mx[k], p[k] + 1u);
                }
            }
        }
    // encode_leaf: palette in order of first appearance (voxel-bit order), then indices.
    uint32_t n = 0, pal0 = 0, pending = 0;
    uint32_t *wslot = words + ((size_t)chunk * 4096u + l1 * 64u + lane) * LEAF_WORDS;
    uint64_t done = ~m;
    while (done != ~0ull) {
        const uint32_t first = (uint32_t)__ffsll((long long)~done) - 1u;
        const uint32_t id = ids[first * 64u + lane];
        for (uint32_t v = first; v < 64u; v++) {
            if (!((done >> v) & 1ull) && ids[v * 64u + lane] == id) {
                done |= 1ull << v;
                idx[v * 64u + lane] = (uint8_t)n;
            }
        }
        if (n == 0) pal0 = id;
        // palette words: pairs lo | hi << 16 from word 1 on
        if (n & 1u) wslot[1u + n / 2u] = pending | (id << 16);
        else pending = id;
        n++;
    }
    uint32_t nwords = 0, entry = 0;
    if (n <= 1u) {
        entry = UNIFORM_BIT | pal0;
    } else {
        if (n & 1u) wslot[1u + n / 2u] = pending;  // odd count: last pair's high half is 0
        const uint32_t bits = n <= 2u ? 1u : (n <= 4u ? 2u : (n <= 16u ? 4u : 8u));
        wslot[0] = n | (bits << 8);
        const uint32_t base = 1u + (n + 1u) / 2u, iw = 64u * bits / 32u;
        for (uint32_t k = 0; k < iw; k++) {
            uint32_t w = 0;
            const uint32_t per = 32u / bits;  // indices per word
            for (uint32_t j = 0; j < per; j++) {
                const uint32_t v = k * per + j;
                if ((m >> v) & 1ull) w |= (uint32_t)idx[v * 64u + lane] << (j * bits);
            }
            wslot[base + k] = w;
        }
        nwords = base + iw;
    }
    // L1 summary (build_local's per-L1 bookkeeping, src/voxel/tree.rs:150-186)
    const uint64_t l1mask = ballot(m != 0ull);
    const bool full = l1mask == ~0ull && ballot(m != ~0ull) == 0ull;
    const bool l1_dry = ballot(dry) != 0ull;
    const uint32_t wl = wave_max(dry ? maxy : 0u);
    const uint32_t lofs = wave_excl_scan(nwords, lane);
    const uint32_t pal_words = wave_sum(nwords);
    const uint32_t solid = wave_sum((uint32_t)__popcll(m));
    const uint32_t em = wave_sum(emitters);
    const uint32_t fl = (ballot(water) ? VX_LT_WATER : 0u) | (ballot(foliage) ? VX_LT_FOLIAGE : 0u) |
                        (ballot(cutout) ? VX_LT_CUTOUT : 0u) | (ballot(glass) ? VX_LT_GLASS : 0u) |
                        (em ? VX_LT_EMITTER : 0u);
    const bool present = m != 0ull;
    uint32_t uni_state = 0, uni_id = 0;
    if (l1mask) {
        const uint32_t lead = (uint32_t)__ffsll((long long)l1mask) - 1u;
        uni_id = (uint32_t)__shfl((int)pal0, (int)lead);
        const bool mixed = ballot(present && (n != 1u || pal0 != uni_id)) != 0ull;
        uni_state = mixed ? 2u : 1u;
    }
    uint32_t cmin[3], cmax[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        cmin[k] = wave_min(present ? mn[k] : 64u);
        cmax[k] = wave_max(present ? mx[k] : 0u);
    }
    LeafInfo li;
    li.mask = m;
    li.nwords = nwords;
    li.entry = n <= 1u ? entry : 0u;
    li.lofs = lofs;
    li.pad = 0;
    leaves[(size_t)chunk * 4096u + l1 * 64u + lane] = li;
    if (lane == 0) {
        L1Info s;
        s.mask = l1mask;
        s.full = full ? 1u : 0u;
        s.dry = l1_dry ? 1u : 0u;
        s.water_line = l1_dry ? wl : 15u;
        s.flags = fl;
        s.emitters = em;
        s.pal_words = pal_words;
        s.solid = solid;
        s.uni_state = uni_state;
        s.uni_id = uni_id;
        s.bmin = cmin[0] | (cmin[1] << 8) | (cmin[2] << 16);
        s.bmax = cmax[0] | (cmax[1] << 8) | (cmax[2] << 16);
        l1s[(size_t)chunk * 64u + l1] = s;
    }
}

__global__ __launch_bounds__(64) void k_local_assemble(const LeafInfo *__restrict__ leaves,
                                                       const uint32_t *__restrict__ words,
                                                       const L1Info *__restrict__ l1s, uint8_t *__restrict__ out,
                                                       size_t slot_bytes) {
    const uint32_t chunk = blockIdx.x, lane = threadIdx.x;
    const L1Info s = l1s[(size_t)chunk * 64u + lane];
    const bool present = s.mask != 0ull;
    const uint32_t cnt = (uint32_t)__popcll(s.mask);
    const uint64_t pres = ballot(present);
    const uint32_t n_l1 = (uint32_t)__popcll(pres);
    const uint32_t rank = (uint32_t)__popcll(pres & ((1ull << lane) - 1ull));
    const uint32_t kept = present && !s.full ? cnt : 0u;
    const uint32_t leaf_start = wave_excl_scan(kept, lane);
    const uint32_t leaf_prefix = wave_excl_scan(cnt, lane);
    const uint32_t pal_base = wave_excl_scan(s.pal_words, lane);
    const uint32_t n_leaves = wave_sum(kept), leaf_count = wave_sum(cnt), n_palette = wave_sum(s.pal_words);
    const uint32_t solid = wave_sum(s.solid), emitters = wave_sum(s.emitters);
    uint32_t fl = 0;
    for (uint32_t f = VX_LT_WATER; f <= VX_LT_EMITTER; f <<= 1)
        if (ballot(present && (s.flags & f))) fl |= f;
    // one block id throughout the chunk: build_local's have_first && all_same
    bool uniform = false;
    uint32_t uid = 0;
    if (pres) {
        const uint32_t lead = (uint32_t)__ffsll((long long)pres) - 1u;
        uid = (uint32_t)__shfl((int)s.uni_id, (int)lead);
        uniform = ballot(present && (s.uni_state != 1u || s.uni_id != uid)) == 0ull;
    }
    if (uniform) fl |= VX_LT_UNIFORM;
    const uint32_t n_attrs = uniform ? 0u : leaf_count, n_pal = uniform ? 0u : n_palette;
    uint32_t bmn[3], bmx[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        bmn[k] = wave_min(present ? (s.bmin >> (8 * k)) & 0xFFu : 64u);
        bmx[k] = wave_max(present ? (s.bmax >> (8 * k)) & 0xFFu : 0u);
    }
    const size_t o_l1 = 64, o_leaves = o_l1 + (size_t)n_l1 * 24u, o_attrs = o_leaves + (size_t)n_leaves * 8u,
                 o_pal = o_attrs + (size_t)n_attrs * 4u;
    const size_t bytes = (o_pal + (size_t)n_pal * 4u + 7u) & ~(size_t)7;
    uint8_t *rec = out + (size_t)chunk * slot_bytes;
    const uint64_t dry_mask = ballot(present && s.dry);  // all lanes: a ballot under lane == 0 sees one lane
    if (lane == 0) {
        vx_local_tree h{};
        h.root_mask = pres;
        h.dry_mask = dry_mask;
        h.n_l1 = n_l1;
        h.n_leaves = n_leaves;
        h.n_attrs = n_attrs;
        h.n_palette = n_pal;
        h.uniform = uniform ? uid : 0u;
        h.flags = fl;
        h.emitters = emitters;
        h.solid_count = solid;
        h.leaf_count = leaf_count;
        if (solid) {
            for (int k = 0; k < 3; k++) {
                h.aabb_min[k] = (uint8_t)bmn[k];
                h.aabb_max[k] = (uint8_t)bmx[k];
            }
        }
        h.bytes = (uint32_t)bytes;
        *(vx_local_tree *)rec = h;
        if (bytes > o_pal + (size_t)n_pal * 4u) *(uint32_t *)(rec + bytes - 4u) = 0u;  // tail padding
    }
    if (present) {
        vx_local_l1 r{s.mask, leaf_start, leaf_prefix, s.full, s.water_line};
        *(vx_local_l1 *)(rec + o_l1 + (size_t)rank * 24u) = r;
    }
    // leaves, attribute entries and palette words, L1 cell by L1 cell (lanes = leaves)
    uint64_t todo = pres;
    while (todo) {
        const uint32_t c = (uint32_t)__ffsll((long long)todo) - 1u;
        todo &= todo - 1ull;
        const uint64_t cm = ((uint64_t)__builtin_amdgcn_readlane((int)(uint32_t)s.mask, (int)c) & 0xFFFFFFFFull) |
                            ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(s.mask >> 32), (int)c) << 32);
        const bool cfull = __builtin_amdgcn_readlane((int)s.full, (int)c) != 0;
        const uint32_t c_start = (uint32_t)__builtin_amdgcn_readlane((int)leaf_start, (int)c);
        const uint32_t c_prefix = (uint32_t)__builtin_amdgcn_readlane((int)leaf_prefix, (int)c);
        const uint32_t c_pal = (uint32_t)__builtin_amdgcn_readlane((int)pal_base, (int)c);
        if (!((cm >> lane) & 1ull)) continue;
        const uint32_t li = (uint32_t)__popcll(cm & ((1ull << lane) - 1ull));
        const size_t g = (size_t)chunk * 4096u + c * 64u + lane;
        const LeafInfo L = leaves[g];
        if (!cfull) *(uint64_t *)(rec + o_leaves + (size_t)(c_start + li) * 8u) = L.mask;
        if (!uniform) {
            const uint32_t off = c_pal + L.lofs;
            *(uint32_t *)(rec + o_attrs + (size_t)(c_prefix + li) * 4u) = L.nwords ? off : L.entry;
            const uint32_t *w = words + g * LEAF_WORDS;
            for (uint32_t k = 0; k < L.nwords; k++) *(uint32_t *)(rec + o_pal + (size_t)(off + k) * 4u) = w[k];
        }
    }
}

}  // namespace vxd

using namespace vxd;
using vxh::ensure;

namespace {

constexpr uint32_t GROUP = 256;  // chunks per device pass

int build_group(vx_ctx *c, const uint16_t *d_dense, uint32_t n, uint8_t *d_out, size_t slot_bytes) {
    int rc;
    if ((rc = ensure(c->build_leaf, (size_t)n * 4096u * sizeof(LeafInfo))) ||
        (rc = ensure(c->build_words, (size_t)n * 4096u * LEAF_WORDS * 4u)) ||
        (rc = ensure(c->build_l1, (size_t)n * 64u * sizeof(L1Info))))
        return rc;
    k_local_leaves<<<n * 64u, 64, 0, c->stream>>>(d_dense, (LeafInfo *)c->build_leaf.p, (uint32_t *)c->build_words.p,
                                                  (L1Info *)c->build_l1.p);
    k_local_assemble<<<n, 64, 0, c->stream>>>((const LeafInfo *)c->build_leaf.p, (const uint32_t *)c->build_words.p,
                                              (const L1Info *)c->build_l1.p, d_out, slot_bytes);
    VX_LAUNCHED();
    return VX_OK;
}

}  // namespace

extern "C" {

int vx_build_local_device(vx_ctx *c, const uint16_t *dense, uint32_t n, void *out, size_t slot_bytes) {
    if (!c || (n && (!dense || !out)) || slot_bytes < VX_LOCAL_TREE_MAX || slot_bytes % 8u) return VX_E_ARG;
    if ((uintptr_t)dense % 8u || (uintptr_t)out % 8u) return VX_E_ARG;  // 8-byte row loads and record writes
    VX_HIPX(hipSetDevice(c->device));
    for (uint32_t b = 0; b < n; b += GROUP) {
        const uint32_t g = std::min(GROUP, n - b);
        int rc = build_group(c, dense + (size_t)b * VOL, g, (uint8_t *)out + (size_t)b * slot_bytes, slot_bytes);
        if (rc) return rc;
    }
    VX_HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_dev_alloc(vx_ctx *c, size_t bytes, void **out) {
    if (!c || !out || !bytes) return VX_E_ARG;
    *out = nullptr;
    VX_HIPX(hipSetDevice(c->device));
    if (hipMalloc(out, bytes) != hipSuccess) {
        *out = nullptr;
        return VX_E_OOM;
    }
    return VX_OK;
}

int vx_dev_free(vx_ctx *c, void *p) {
    if (!c) return VX_E_ARG;
    if (!p) return VX_OK;
    VX_HIPX(hipSetDevice(c->device));
    VX_HIPX(hipStreamSynchronize(c->stream));
    VX_HIPX(hipFree(p));
    return VX_OK;
}

int vx_dev_upload(vx_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) return VX_E_ARG;
    VX_HIPX(hipSetDevice(c->device));
    VX_HIPX(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    VX_HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_dev_download(vx_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) return VX_E_ARG;
    VX_HIPX(hipSetDevice(c->device));
    VX_HIPX(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    VX_HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_build_local(vx_ctx *c, const uint16_t *dense, uint32_t n, void *out, size_t slot_bytes) {
    if (!c || (n && (!dense || !out)) || slot_bytes < VX_LOCAL_TREE_MAX || slot_bytes % 8u) return VX_E_ARG;
    VX_HIPX(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    for (uint32_t b = 0; b < n; b += GROUP) {
        const uint32_t g = std::min(GROUP, n - b);
        int rc;
        if ((rc = ensure(c->build_dense, (size_t)g * VOL * 2u)) ||
            (rc = ensure(c->build_out, (size_t)g * VX_LOCAL_TREE_MAX)))
            return rc;
        VX_HIPX(hipMemcpyAsync(c->build_dense.p, dense + (size_t)b * VOL, (size_t)g * VOL * 2u,
                               hipMemcpyHostToDevice, st));
        if ((rc = build_group(c, (const uint16_t *)c->build_dense.p, g, (uint8_t *)c->build_out.p,
                              VX_LOCAL_TREE_MAX)))
            return rc;
        // the headers first, then each record's own bytes
        std::vector<vx_local_tree> hdr(g);
        VX_HIPX(hipMemcpy2DAsync(hdr.data(), sizeof(vx_local_tree), c->build_out.p, VX_LOCAL_TREE_MAX,
                                 sizeof(vx_local_tree), g, hipMemcpyDeviceToHost, st));
        VX_HIPX(hipStreamSynchronize(st));
        for (uint32_t i = 0; i < g; i++) {
            if (hdr[i].bytes < sizeof(vx_local_tree) || hdr[i].bytes > VX_LOCAL_TREE_MAX) return VX_E_HIP;
            VX_HIPX(hipMemcpyAsync((uint8_t *)out + (size_t)(b + i) * slot_bytes,
                                   (const uint8_t *)c->build_out.p + (size_t)i * VX_LOCAL_TREE_MAX, hdr[i].bytes,
                                   hipMemcpyDeviceToHost, st));
        }
        VX_HIPX(hipStreamSynchronize(st));
    }
    return VX_OK;
}

}  // extern "C"
