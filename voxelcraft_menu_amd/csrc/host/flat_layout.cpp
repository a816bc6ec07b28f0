// flat_layout.cpp -- the three pools re-laid out as one block per chunk.
//
// The reference keeps three shared pools (geometry.rs:12-52, attributes.rs:6-88): L1
// runs and leaf runs are interned, so identical runs are stored once and chunks point
// into them wherever their runs happen to lie. The march reads, per chunk, one L1 node
// per L1 cell change, one leaf per leaf change and attribute words where it needs them
// (common.wgsl:697-760). This layout gives each chunk one contiguous block that starts
// on a 128-byte line:
//
//   [L1 nodes of the root run][leaf runs of its L1 cells, in run order][attribute table]
//   [palettes the table points to, each once]
//
// with every pointer rewritten into the block: the root's run_ptr (16-byte units), each
// L1 node's leaf run pointer (8-byte units), the chunk's attr_base and the table's
// palette offsets (4-byte units). Element indexes of the three kinds address the same
// buffer, so one allocation serves as all three pools. Masks, prefixes, water lines,
// dry masks and block ids are unchanged, so every lookup of common.wgsl:266-356 and of
// the march returns what it returns on the reference layout. Runs shared by several
// chunks are copied into each of them.
#include <cstring>
#include <unordered_map>
#include <vector>

#include "vxhost.h"

namespace {

constexpr uint32_t FULL = 0x80000000u, PTR = 0x7FFFFFFFu, UNIFORM = 0x80000000u;
constexpr uint32_t ATTR_UNIFORM_FLAG = 1u;
constexpr size_t LINE_WORDS = 32;  // 128-byte blocks

uint32_t popc64(uint64_t m) { return (uint32_t)__builtin_popcountll(m); }

struct Src {
    const uint32_t *in, *lf, *br;
    size_t n_in, n_lf, n_br;  // elements: 16-byte nodes, 8-byte leaves, words
};

// Appends one chunk's block at word offset `at` of out (out may be null: size only).
// Returns the offset after the block, or 0 on a pointer outside its pool.
size_t emit_chunk(const Src &s, const vx_gpu_chunk &c, vx_gpu_chunk &o, uint32_t *out, size_t at) {
    o = c;
    const uint64_t root_mask = (uint64_t)c.root[0] | ((uint64_t)c.root[1] << 32);
    const bool root_full = (c.root[2] & FULL) != 0u;
    const bool uniform = (c.attr_flags & ATTR_UNIFORM_FLAG) != 0u;
    size_t w = at;
    // L1 nodes: the root run (a full root has none, the march synthesises them).
    struct L1 { uint32_t node[4]; };
    std::vector<L1> l1;
    if (!root_full && root_mask) {
        const size_t base = c.root[2] & PTR, k = popc64(root_mask);
        if (base + k > s.n_in) return 0;
        l1.resize(k);
        std::memcpy(l1.data(), s.in + base * 4, k * 16);
        o.root[2] = (uint32_t)(w / 4);
        w += 4 * k;
    } else if (!root_full) {
        o.root[2] = 0;
    }
    // Leaf runs of the non-full L1 cells.
    for (L1 &n : l1) {
        if (n.node[2] & FULL) continue;
        const uint64_t m = (uint64_t)n.node[0] | ((uint64_t)n.node[1] << 32);
        const size_t base = n.node[2] & PTR, k = popc64(m);
        if (base + k > s.n_lf) return 0;
        if (out) std::memcpy(out + w, s.lf + base * 2, k * 8);
        n.node[2] = (uint32_t)(w / 2);
        w += 2 * k;
    }
    if (out && !l1.empty()) std::memcpy(out + (o.root[2] * (size_t)4), l1.data(), l1.size() * 16);
    // Attribute table: entries up to the largest prefix + leaf count of an L1 cell (64 per
    // full cell; a full root synthesises cell l1b at prefix l1b * 64).
    if (!uniform) {
        size_t len = 0;
        if (root_full) {
            len = 64 * 64;
        } else {
            for (const L1 &n : l1) {
                const uint64_t m = (uint64_t)n.node[0] | ((uint64_t)n.node[1] << 32);
                const size_t cnt = (n.node[2] & FULL) ? 64 : popc64(m);
                const size_t end = (size_t)(n.node[3] & 0x0FFFFFFFu) + cnt;
                if (end > len) len = end;
            }
        }
        const size_t tbase = c.attr_base;
        if (tbase + len > s.n_br) return 0;
        const size_t t = w;
        o.attr_base = (uint32_t)t;
        w += len;
        std::unordered_map<uint32_t, uint32_t> moved;  // palette offset -> new offset
        for (size_t i = 0; i < len; i++) {
            uint32_t e = s.br[tbase + i];
            if (!(e & UNIFORM)) {
                auto it = moved.find(e);
                if (it == moved.end()) {
                    if (e >= s.n_br) return 0;
                    const uint32_t hdr = s.br[e];
                    const size_t n_pal = hdr & 0xFFu, bits = (hdr >> 8) & 0xFFu;
                    const size_t plen = 1 + (n_pal + 1) / 2 + 2 * bits;  // header, ids, 64 indexes
                    if (e + plen > s.n_br) return 0;
                    if (out) std::memcpy(out + w, s.br + e, plen * 4);
                    it = moved.emplace(e, (uint32_t)w).first;
                    w += plen;
                }
                e = it->second;
            }
            if (out) out[t + i] = e;
        }
    }
    if (w == at) w += 4;  // keep blocks distinct (and the buffer non-empty)
    return (w + LINE_WORDS - 1) / LINE_WORDS * LINE_WORDS;
}

}  // namespace

extern "C" {

int vxw_flat_layout(const uint32_t *inners, size_t n_inners, const uint32_t *leaves, size_t n_leaves,
                    const uint32_t *bricks, size_t n_bricks, const vx_gpu_chunk *chunks, uint32_t n,
                    vx_gpu_chunk *chunks_out, uint32_t *flat, size_t cap_words, size_t *words) {
    if (!words || (n && (!chunks || !chunks_out))) return VXW_E_ARG;
    if ((n_inners && !inners) || (n_leaves && !leaves) || (n_bricks && !bricks)) return VXW_E_ARG;
    const Src s{inners, leaves, bricks, n_inners, n_leaves, n_bricks};
    // Pass 1 sizes the buffer, pass 2 fills it (the same walk, so offsets agree).
    size_t w = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (!(chunks[i].light_flags & 1u)) return VXW_E_ARG;  // light tables are not moved (uniform light only)
        vx_gpu_chunk tmp;
        const size_t next = emit_chunk(s, chunks[i], tmp, nullptr, w);
        if (!next) return VXW_E_ARG;
        w = next;
    }
    *words = w;
    if (!flat) return VXW_OK;
    if (cap_words < w) return VXW_E_ARG;
    std::memset(flat, 0, w * 4);
    size_t at = 0;
    for (uint32_t i = 0; i < n; i++) at = emit_chunk(s, chunks[i], chunks_out[i], flat, at);
    return VXW_OK;
}

}  // extern "C"
