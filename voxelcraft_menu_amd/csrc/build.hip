// build.hip -- build_local (src/voxel/tree.rs:89-204) on the device, many chunks at once.
//
// Two passes over a batch of dense 64^3 chunks:
//   k_local_leaves    one 64-lane wave per L1 cell (16^3 voxels), one lane per leaf (4^3):
//                     the leaf's occupancy mask, its palette encoding (encode_leaf,
//                     src/voxel/attributes.rs:90-166) into a fixed 49-word slot, and the
//                     L1 summary (mask, full, dry, water line, flags, counts, bounds) by
//                     wave reductions. A block is the 4 L1 cells of one row along x: a
//                     row of the dense chunk is 128 bytes, a quarter per L1 cell, so the
//                     block reads each line once, on one CU (one wave per L1 cell as
//                     separate blocks lands the four quarters on different XCDs, whose L2s
//                     each fetch the whole line);
//   k_local_assemble  per chunk, one lane per L1 cell: prefix sums over the L1 cells give
//                     leaf_start, leaf_prefix and the palette offsets, and the chunk's flat
//                     vx_local_tree record is written (header, L1 nodes, kept leaf masks,
//                     attribute entries, palette words); four waves split the L1 cells.
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

// Bit spreading for the index words: the low 16 / 8 / 4 bits of x to every 2nd / 4th /
// 8th bit.
__device__ __forceinline__ uint32_t spread_2(uint64_t v) {
    uint32_t x = (uint32_t)v & 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    return (x | (x << 1)) & 0x55555555u;
}
__device__ __forceinline__ uint32_t spread_4(uint64_t v) {
    uint32_t x = (uint32_t)v & 0xFFu;
    x = (x | (x << 12)) & 0x000F000Fu;
    x = (x | (x << 6)) & 0x03030303u;
    return (x | (x << 3)) & 0x11111111u;
}
__device__ __forceinline__ uint32_t spread_8(uint64_t v) {
    uint32_t x = (uint32_t)v & 0xFu;
    x = (x | (x << 14)) & 0x00030003u;
    return (x | (x << 7)) & 0x01010101u;
}

constexpr uint32_t LEAVES_BLOCK = 256;  // 4 waves: the L1 cells l1x = 0..3 of one row

__global__ __launch_bounds__(LEAVES_BLOCK) void k_local_leaves(const uint16_t *__restrict__ dense,
                                                               LeafInfo *__restrict__ leaves,
                                                               uint32_t *__restrict__ words,
                                                               L1Info *__restrict__ l1s) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t chunk = blockIdx.x >> 4, l1 = wave | ((blockIdx.x & 15u) << 2);
    const uint32_t l1x = l1 & 3u, l1y = (l1 >> 2) & 3u, l1z = l1 >> 4;
    const uint32_t lx = lane & 3u, ly = (lane >> 2) & 3u, lz = lane >> 4;
    const uint32_t bx = l1x * 16u + lx * 4u, by = l1y * 16u + ly * 4u, bz = l1z * 16u + lz * 4u;
    const uint16_t *src = dense + (size_t)chunk * VOL;
    uint2 rows[16];  // all 16 row loads in flight before any is used
#pragma unroll
    for (uint32_t r = 0; r < 16u; r++)
        rows[r] = *(const uint2 *)(src + (bx | ((bz + (r >> 2)) << 6) | ((by + (r & 3u)) << 12)));
    // One branch-free pass in registers, voxel-bit order (x fastest), masks kept as 32-bit
    // halves: occupancy, dry (non-water) and emitter voxels, one bit per block-table row
    // present (ids past the table count as its last row) and the largest id.
    uint32_t m_h[2] = {0u, 0u}, dry_h[2] = {0u, 0u}, em_h[2] = {0u, 0u};
    uint32_t idbits = 0, idmax = 0;
#pragma unroll
    for (uint32_t r = 0; r < 16u; r++) {
        const uint32_t w4[4] = {rows[r].x & 0xFFFFu, rows[r].x >> 16, rows[r].y & 0xFFFFu, rows[r].y >> 16};
#pragma unroll
        for (uint32_t vx = 0; vx < 4u; vx++) {
            const uint32_t id = w4[vx], vb = vx | (r << 2), h = vb >> 5, sh = vb & 31u;  // r = vz * 4 + vy
            const uint32_t cid = min(id, 25u);
            m_h[h] |= min(id, 1u) << sh;
            dry_h[h] |= min(id ^ (uint32_t)B_WATER, 1u) << sh;
            em_h[h] |= ((EMITTER_IDS >> cid) & 1u) << sh;
            idbits |= 1u << cid;
            idmax = max(idmax, id);
        }
        // accumulate row by row: left free, the compiler extracts all 64 ids first and
        // sums them as trees, which holds some 200 registers
        asm volatile("" : "+v"(m_h[0]), "+v"(m_h[1]), "+v"(dry_h[0]), "+v"(dry_h[1]), "+v"(em_h[0]), "+v"(em_h[1]),
                     "+v"(idbits), "+v"(idmax));
    }
    const uint64_t m = ((uint64_t)m_h[1] << 32) | m_h[0];
    const uint64_t dry_m = (((uint64_t)dry_h[1] << 32) | dry_h[0]) & m;
    const uint64_t em_m = ((uint64_t)em_h[1] << 32) | em_h[0];
    idbits &= ~1u;  // air
    // Several table rows means several ids. A leaf whose ids are all at or past the last
    // row is checked voxel by voxel against its largest id.
    bool mixed = (idbits & (idbits - 1u)) != 0u;
    const bool past_only = idbits == (1u << 25);
    const uint32_t pal0 = past_only ? idmax : (idbits ? (uint32_t)__builtin_ctz(idbits) : 0u);
    // The later passes re-extract ids from the row registers: keeping the 64 extracted ids
    // of the pass above alive across them would cost 64 registers.
#pragma unroll
    for (uint32_t r = 0; r < 16u; r++) asm volatile("" : "+v"(rows[r].x), "+v"(rows[r].y));
    if (past_only) {
#pragma unroll
        for (uint32_t r = 0; r < 16u; r++) {
            const uint32_t w4[4] = {rows[r].x & 0xFFFFu, rows[r].x >> 16, rows[r].y & 0xFFFFu, rows[r].y >> 16};
#pragma unroll
            for (uint32_t vx = 0; vx < 4u; vx++) mixed = mixed || (w4[vx] != 0u && w4[vx] != idmax);
        }
    }
    const bool water = (idbits >> B_WATER) & 1u, glass = (idbits >> B_GLASS) & 1u;
    const bool foliage = (idbits & FOLIAGE_IDS) != 0u, cutout = (idbits & CUTOUT_IDS) != 0u;
    const bool dry = dry_m != 0ull;
    const uint32_t emitters = (uint32_t)__popcll(em_m);
    // Bounds and water line from the masks (bit x | y << 2 | z << 4): fold to 4-bit
    // occupancies per axis.
    auto axis_bits = [](uint64_t v, int axis) {
        if (axis == 0) {  // x: fold z (16-bit groups), then y (4-bit groups)
            uint32_t t = (uint32_t)(v | (v >> 32));
            t |= t >> 16;
            t |= t >> 8;
            t |= t >> 4;
            return t & 0xFu;
        }
        if (axis == 1) {  // y: fold z, then x within each 4-bit group
            uint32_t t = (uint32_t)(v | (v >> 32));
            t |= t >> 16;
            t = (t | (t >> 1) | (t >> 2) | (t >> 3)) & 0x1111u;
            return (t & 1u) | ((t >> 3) & 2u) | ((t >> 6) & 4u) | ((t >> 9) & 8u);
        }
        // z: fold x and y within each 16-bit group
        uint64_t t = v | (v >> 1) | (v >> 2) | (v >> 3);
        t |= (t >> 4) | (t >> 8) | (t >> 12);
        t &= 0x0001000100010001ull;
        return (uint32_t)((t & 1u) | ((t >> 15) & 2u) | ((t >> 30) & 4u) | ((t >> 45) & 8u));
    };
    uint32_t mn[3] = {64u, 64u, 64u}, mx[3] = {0u, 0u, 0u};
    if (m) {
        const uint32_t base[3] = {bx, by, bz};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const uint32_t ab = axis_bits(m, k);
            mn[k] = base[k] + (uint32_t)__builtin_ctz(ab);
            mx[k] = base[k] + 32u - (uint32_t)__builtin_clz(ab);
        }
    }
    const uint32_t maxy = dry ? ly * 4u + 31u - (uint32_t)__builtin_clz(axis_bits(dry_m, 1)) : 0u;
    uint32_t n = m ? 1u : 0u;
    uint32_t *wslot = words + ((size_t)chunk * 4096u + l1 * 64u + lane) * LEAF_WORDS;
    // encode_leaf for a leaf of several ids: the palette in order of first appearance
    // (voxel-bit order), each entry's voxels found by one compare pass over the row
    // registers; index bit b of every voxel is kept as a 64-bit plane (n <= 64 needs 6).
    uint64_t planes[6] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
    if (mixed) {
        n = 0;
        uint32_t pending = 0;
        uint64_t todo = m;
        while (todo) {
#pragma unroll
            for (uint32_t r = 0; r < 16u; r++) asm volatile("" : "+v"(rows[r].x), "+v"(rows[r].y));
            const uint32_t first = (uint32_t)__ffsll((long long)todo) - 1u;
            const uint32_t fr = first >> 2;
            uint2 w = rows[0];
#pragma unroll
            for (uint32_t r = 1; r < 16u; r++) w = fr == r ? rows[r] : w;
            const uint32_t half = (first & 2u) ? w.y : w.x;
            const uint32_t id = (first & 1u) ? half >> 16 : half & 0xFFFFu;
            uint32_t eq_lo = 0, eq_hi = 0;
#pragma unroll
            for (uint32_t r = 0; r < 16u; r++) {
                const uint32_t w4[4] = {rows[r].x & 0xFFFFu, rows[r].x >> 16, rows[r].y & 0xFFFFu, rows[r].y >> 16};
#pragma unroll
                for (uint32_t vx = 0; vx < 4u; vx++) {
                    const uint32_t vb = vx | (r << 2);  // r = vz * 4 + vy
                    if (vb < 32u) eq_lo |= w4[vx] == id ? 1u << vb : 0u;
                    else eq_hi |= w4[vx] == id ? 1u << (vb - 32u) : 0u;
                }
                asm volatile("" : "+v"(eq_lo), "+v"(eq_hi));  // row by row, as in the pass above
            }
            const uint64_t eq = (((uint64_t)eq_hi << 32) | eq_lo) & todo;
            todo &= ~eq;
#pragma unroll
            for (uint32_t b2 = 0; b2 < 6u; b2++) planes[b2] |= ((n >> b2) & 1u) ? eq : 0ull;
            // palette words: pairs lo | hi << 16 from word 1 on
            if (n & 1u) wslot[1u + n / 2u] = pending | (id << 16);
            else pending = id;
            n++;
        }
        if (n & 1u) wslot[1u + n / 2u] = pending;  // odd count: last pair's high half is 0
    }
    uint32_t nwords = 0, entry = 0;
    if (n <= 1u) {
        entry = UNIFORM_BIT | pal0;
    } else {
        // index words: voxel v's index at bit (v % (32 / bits)) * bits of word v / (32 / bits)
        const uint32_t bits = n <= 2u ? 1u : (n <= 4u ? 2u : (n <= 16u ? 4u : 8u));
        wslot[0] = n | (bits << 8);
        const uint32_t base = 1u + (n + 1u) / 2u, iw = 64u * bits / 32u;
        uint32_t *iwp = wslot + base;
        if (bits == 1u) {
            iwp[0] = (uint32_t)planes[0];
            iwp[1] = (uint32_t)(planes[0] >> 32);
        } else if (bits == 2u) {
#pragma unroll
            for (uint32_t k = 0; k < 4u; k++)
                iwp[k] = spread_2(planes[0] >> (16u * k)) | (spread_2(planes[1] >> (16u * k)) << 1);
        } else if (bits == 4u) {
#pragma unroll
            for (uint32_t k = 0; k < 8u; k++) {
                uint32_t wd = 0;
#pragma unroll
                for (uint32_t b2 = 0; b2 < 4u; b2++) wd |= spread_4(planes[b2] >> (8u * k)) << b2;
                iwp[k] = wd;
            }
        } else {
#pragma unroll
            for (uint32_t k = 0; k < 16u; k++) {
                uint32_t wd = 0;
#pragma unroll
                for (uint32_t b2 = 0; b2 < 6u; b2++) wd |= spread_8(planes[b2] >> (4u * k)) << b2;
                iwp[k] = wd;
            }
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
        const bool l1_mixed = ballot(present && (mixed || pal0 != uni_id)) != 0ull;
        uni_state = l1_mixed ? 2u : 1u;
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

constexpr uint32_t ASSEMBLE_WAVES = 8, ASSEMBLE_BLOCK = 64u * ASSEMBLE_WAVES;

__global__ __launch_bounds__(ASSEMBLE_BLOCK) void k_local_assemble(const LeafInfo *__restrict__ leaves,
                                                                   const uint32_t *__restrict__ words,
                                                                   const L1Info *__restrict__ l1s,
                                                                   uint8_t *__restrict__ out, size_t slot_bytes) {
    // Every wave computes the chunk's scans (64 L1 summaries); wave 0 writes the header
    // and the L1 nodes, and the waves take the L1 cells' leaves in turn (c % 8 == wave).
    const uint32_t chunk = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
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
    if (wave == 0 && lane == 0) {
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
    if (wave == 0 && present) {
        vx_local_l1 r{s.mask, leaf_start, leaf_prefix, s.full, s.water_line};
        *(vx_local_l1 *)(rec + o_l1 + (size_t)rank * 24u) = r;
    }
    // leaves, attribute entries and palette words, L1 cell by L1 cell (lanes = leaves)
    uint64_t todo = pres & (0x0101010101010101ull << wave);
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

// Chunks per device pass. The pass's scratch is 800 KB of palette slots and 100 KB of
// leaf and L1 summaries per chunk (0.9 GB at 1024 chunks).
constexpr uint32_t GROUP = 1024;
// Chunks per pass of the host-buffer entry point, which also stages the dense chunks and
// the records in device memory (512 KB + 834 KB per chunk).
constexpr uint32_t HOST_GROUP = 256;

int build_group(vx_ctx *c, const uint16_t *d_dense, uint32_t n, uint8_t *d_out, size_t slot_bytes) {
    int rc;
    if ((rc = ensure(c->build_leaf, (size_t)n * 4096u * sizeof(LeafInfo))) ||
        (rc = ensure(c->build_words, (size_t)n * 4096u * LEAF_WORDS * 4u)) ||
        (rc = ensure(c->build_l1, (size_t)n * 64u * sizeof(L1Info))))
        return rc;
    k_local_leaves<<<n * 16u, LEAVES_BLOCK, 0, c->stream>>>(d_dense, (LeafInfo *)c->build_leaf.p,
                                                             (uint32_t *)c->build_words.p, (L1Info *)c->build_l1.p);
    k_local_assemble<<<n, ASSEMBLE_BLOCK, 0, c->stream>>>((const LeafInfo *)c->build_leaf.p,
                                                          (const uint32_t *)c->build_words.p,
                                                          (const L1Info *)c->build_l1.p, d_out, slot_bytes);
    VX_LAUNCHED();
    return VX_OK;
}

}  // namespace

int vxh::build_local_async(vx_ctx *c, const uint16_t *d_dense, uint32_t n, uint8_t *d_out, size_t slot_bytes) {
    for (uint32_t b = 0; b < n; b += GROUP) {
        const uint32_t g = std::min(GROUP, n - b);
        int rc = build_group(c, d_dense + (size_t)b * VOL, g, d_out + (size_t)b * slot_bytes, slot_bytes);
        if (rc) return rc;
    }
    return VX_OK;
}

extern "C" {

int vx_build_local_device(vx_ctx *c, const uint16_t *dense, uint32_t n, void *out, size_t slot_bytes) {
    if (!c || (n && (!dense || !out)) || slot_bytes < VX_LOCAL_TREE_MAX || slot_bytes % 8u) return VX_E_ARG;
    if ((uintptr_t)dense % 8u || (uintptr_t)out % 8u) return VX_E_ARG;  // 8-byte row loads and record writes
    VX_HIPX(hipSetDevice(c->device));
    int rc = vxh::build_local_async(c, dense, n, (uint8_t *)out, slot_bytes);
    if (rc) return rc;
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
    for (uint32_t b = 0; b < n; b += HOST_GROUP) {
        const uint32_t g = std::min(HOST_GROUP, n - b);
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
