// dworld.hip -- a device-resident world: the pools in the per-chunk block layout
// (vxw_flat_layout: one 128-byte aligned block per chunk holding its L1 nodes, leaf runs,
// attribute table and palettes) and the table of chunk records, kept on the GPU, with
// edit batches applied there.
//
// An edit batch (journal records, src/journal.rs:116-122) is applied as World::set_block
// applies it record by record (src/voxel/world.rs:288-529), with the same results for
// rendering:
//   1. k_dw_decode: every chunk an edit falls in is decoded to dense ids (one thread per
//      leaf: its mask, attribute entry and palette);
//   2. k_dw_edits: one thread per edited voxel replays that voxel's edits in order; an
//      edit is effective when its id differs from the voxel's current one (set_block's
//      two early returns). Effective edits mark their L1 cell touched, grow the dry mask
//      with non-water, non-air ids, add the water / foliage / cutout / glass flags of
//      their ids, and grow the box by voxels that become solid -- the record fields
//      set_block only ever grows;
//   3. build_local of the edited chunks on the device (k_local_leaves / k_local_assemble);
//   4. k_dw_place writes each chunk's new block at the end of the used space and its new
//      record: L1 cells touched by an effective edit get water line 0 (15 when full), as
//      set_block's rebuilt node carries, the others keep their old node's water line.
// Blocks are appended: the space of a replaced block is not reused. Keys and shadow words
// rendered from this world equal those of the host world the same edits were applied to
// (the pools' bytes differ: no interning, no free lists).
#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "ctx.hpp"

namespace vxd {

constexpr uint32_t DW_VOL = 64u * 64u * 64u;
constexpr uint16_t DW_AIR = 0, DW_WATER = 13, DW_GLASS = 10;
constexpr uint32_t DW_FOLIAGE_IDS = (1u << 18) | (1u << 24) | (1u << 25), DW_CUTOUT_IDS = (1u << 7) | (1u << 16);

// Per edited chunk of a batch.
struct DwSlot {
    uint32_t chunk;                    // table index
    uint32_t flags;                    // ATTR_HAS_* of the ids effective edits placed
    unsigned long long touched, dry;   // L1 cells with effective edits / with non-water placements
    int32_t bmin[3], bmax[3];          // voxels that became solid: min and max + 1
    uint32_t applied;                  // effective edits
    uint32_t words, offset;            // the new block's size and place (words)
    uint32_t pad;
};

__device__ __forceinline__ uint32_t dense_index(uint32_t x, uint32_t y, uint32_t z) { return x | (z << 6) | (y << 12); }

// Dense ids of the chunks in the slots. One block per (slot, 64 x 16 x 16 slab of four L1
// cells along x): the slab's 256 leaves (mask, attribute entry) are read into LDS, then
// each wave writes four whole 128-byte z rows per store (8 bytes = one leaf row per lane).
__global__ __launch_bounds__(256) void k_dw_decode(Pools p, const DwSlot *__restrict__ slots,
                                                   uint16_t *__restrict__ dense) {
    __shared__ unsigned long long s_mask[256];
    __shared__ uint32_t s_entry[256];
    const uint32_t s = blockIdx.x >> 4, l1y = blockIdx.x & 3u, l1z = (blockIdx.x >> 2) & 3u, t = threadIdx.x;
    const ChunkK c = ld_chunk(p, slots[s].chunk);
    const bool uniform = (c.attr_flags & ATTR_UNIFORM) != 0u;
    {
        const uint32_t l1b = (t >> 6) | (l1y << 2) | (l1z << 4), lfb = t & 63u;
        const uint64_t root_mask = mk64(c.root[0], c.root[1]);
        uint64_t leaf = 0;
        uint32_t entry = 0;
        if (mask_has(root_mask, l1b)) {
            uint4 l1;
            if (c.root[2] & FULL_BIT) l1 = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, FULL_BIT, l1b * 64u);
            else l1 = ld_inner(p, (c.root[2] & PTR_MASK) + mask_below(root_mask, l1b));
            const uint64_t l1m = mk64(l1.x, l1.y);
            if (l1.z & FULL_BIT) {
                leaf = ~0ull;
                if (!uniform) entry = ld_brick(p, c.attr_base + (l1.w & 0x0FFFFFFFu) + lfb);
            } else if (mask_has(l1m, lfb)) {
                const uint32_t li = mask_below(l1m, lfb);
                const uint2 lf = ld_leaf(p, (l1.z & PTR_MASK) + li);
                leaf = mk64(lf.x, lf.y);
                if (!uniform) entry = ld_brick(p, c.attr_base + (l1.w & 0x0FFFFFFFu) + li);
            }
        }
        s_mask[t] = leaf;
        s_entry[t] = entry;
    }
    __syncthreads();
    uint16_t *out = dense + (size_t)s * DW_VOL;
    const uint32_t seg = t & 15u, zz = t >> 4;  // x = 4 seg .. 4 seg + 3, z = 16 l1z + zz
    const uint32_t l1x = seg >> 2, lx = seg & 3u, lz = zz >> 2, vz = zz & 3u;
    for (uint32_t yy = 0; yy < 16u; yy++) {
        const uint32_t ly = yy >> 2, vy = yy & 3u;
        const uint32_t slot = l1x * 64u + (lx | (ly << 2) | (lz << 4));
        const uint32_t bits = (uint32_t)(s_mask[slot] >> (vy * 4u + vz * 16u)) & 15u;
        uint32_t v[4] = {DW_AIR, DW_AIR, DW_AIR, DW_AIR};
        if (bits) {
            const uint32_t entry = s_entry[slot];
#pragma unroll
            for (uint32_t vx = 0; vx < 4u; vx++)
                if ((bits >> vx) & 1u) v[vx] = uniform ? c.attr_base : leaf_block(p, entry, vx, vy, vz);
        }
        *(uint2 *)(out + dense_index(seg * 4u, l1y * 16u + yy, l1z * 16u + zz)) =
            make_uint2(v[0] | (v[1] << 16), v[2] | (v[3] << 16));
    }
}

// One thread per edited voxel: its edits in order (sorted by slot and voxel, stable).
// word = slot << 18 | dense index; ids beside.
__global__ __launch_bounds__(256) void k_dw_edits(const uint32_t *__restrict__ word, const uint16_t *__restrict__ ids,
                                                  const uint32_t *__restrict__ run_start, uint32_t n_runs,
                                                  DwSlot *__restrict__ slots, uint16_t *__restrict__ dense) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_runs) return;
    const uint32_t b = run_start[r], e = run_start[r + 1];
    const uint32_t s = word[b] >> 18, di = word[b] & 0x3FFFFu;
    const uint32_t x = di & 63u, z = (di >> 6) & 63u, y = di >> 12;
    const uint32_t l1b = cell_bit(x >> 4, y >> 4, z >> 4);
    uint16_t *cell = dense + (size_t)s * DW_VOL + di;
    uint32_t cur = *cell;
    uint32_t applied = 0, flags = 0;
    bool touched = false, dry = false, became = false;
    for (uint32_t i = b; i < e; i++) {
        const uint32_t id = ids[i];
        if (id == cur) continue;  // set_block's no-op cases: air to air, the same block
        applied++;
        touched = true;
        if (id != DW_WATER && id != DW_AIR) dry = true;
        if (id == DW_WATER) flags |= ATTR_HAS_WATER;
        if ((DW_FOLIAGE_IDS >> min(id, 25u)) & 1u) flags |= 4u;
        if ((DW_CUTOUT_IDS >> min(id, 25u)) & 1u) flags |= ATTR_HAS_CUTOUT;
        if (id == DW_GLASS) flags |= ATTR_HAS_GLASS;
        if (cur == DW_AIR) became = true;
        cur = id;
    }
    *cell = (uint16_t)cur;
    DwSlot &sl = slots[s];
    if (applied) atomicAdd(&sl.applied, applied);
    if (flags) atomicOr(&sl.flags, flags);
    if (touched) atomicOr(&sl.touched, 1ull << l1b);
    if (dry) atomicOr(&sl.dry, 1ull << l1b);
    if (became) {
        atomicMin(&sl.bmin[0], (int32_t)x); atomicMin(&sl.bmin[1], (int32_t)y); atomicMin(&sl.bmin[2], (int32_t)z);
        atomicMax(&sl.bmax[0], (int32_t)x + 1); atomicMax(&sl.bmax[1], (int32_t)y + 1);
        atomicMax(&sl.bmax[2], (int32_t)z + 1);
    }
}

// Batch state in device memory: the used words of the pools (blocks are appended), the
// capacity, effective edits so far, and whether a pass did not fit.
struct DwState {
    unsigned long long used, cap, effective;
    uint32_t overflow, placed;  // placed: chunks whose new blocks were placed
};

constexpr uint32_t DW_GROUP = 1024;  // chunks per pass: 512 MB of dense ids, 850 MB of records

// The new blocks' sizes (words, a multiple of 32) and places: one thread per slot, a
// prefix sum over the pass, appended at the used end when the pass fits (else the state
// says overflow and the pass places nothing).
__global__ __launch_bounds__(DW_GROUP) void k_dw_offsets(const uint8_t *__restrict__ recs, size_t slot_bytes,
                                                         uint32_t n, DwSlot *__restrict__ slots, DwState *st) {
    __shared__ uint32_t sc[DW_GROUP];
    __shared__ unsigned long long applied_sum;
    const uint32_t t = threadIdx.x;
    uint32_t words = 0;
    if (t == 0) applied_sum = 0;
    if (t < n) {
        const vx_local_tree *h = (const vx_local_tree *)(recs + (size_t)t * slot_bytes);
        const vx_local_l1 *l1 = (const vx_local_l1 *)(h + 1);
        bool all_full = h->n_l1 == 64u;
        for (uint32_t i = 0; all_full && i < h->n_l1; i++) all_full = l1[i].full != 0u;
        const bool uniform = (h->flags & VX_LT_UNIFORM) || h->n_attrs == 0u;
        uint32_t w = (all_full ? 0u : 4u * h->n_l1) + 2u * h->n_leaves;
        if (!uniform) w += h->n_attrs + h->n_palette;
        if (w == 0u) w = 4u;
        words = (w + 31u) & ~31u;
        slots[t].words = words;
    }
    sc[t] = words;
    __syncthreads();
    for (uint32_t d = 1; d < DW_GROUP; d <<= 1) {
        const uint32_t v = t >= d ? sc[t - d] : 0u;
        __syncthreads();
        sc[t] += v;
        __syncthreads();
    }
    if (t < n && slots[t].applied) atomicAdd(&applied_sum, (unsigned long long)slots[t].applied);
    const unsigned long long used = st->used, total = sc[DW_GROUP - 1];
    const bool fits = st->overflow == 0u && used + total <= st->cap;
    if (t < n) slots[t].offset = (uint32_t)(used + sc[t] - words);
    __syncthreads();  // every thread has read the state
    if (t == 0) {
        if (fits) {
            st->used = used + total;
            st->effective += applied_sum;
            st->placed += n;
        } else {
            st->overflow = 1u;
        }
    }
}

// Writes each slot's block at its offset in flat and the chunk's new record, reading the
// old block's L1 nodes from old_flat (the same buffer for an edit, the old one for a
// compaction); one wave per slot.
__global__ __launch_bounds__(64) void k_dw_place(const uint8_t *__restrict__ recs, size_t slot_bytes,
                                                 const DwSlot *__restrict__ slots, const DwState *st,
                                                 const uint32_t *old_flat, uint32_t *flat,
                                                 vx_gpu_chunk *__restrict__ table, uint32_t *__restrict__ solid) {
    if (st->overflow) return;  // the pass did not fit: its chunks keep their old blocks
    const uint32_t s = blockIdx.x, lane = threadIdx.x;
    const DwSlot sl = slots[s];
    const vx_local_tree *h = (const vx_local_tree *)(recs + (size_t)s * slot_bytes);
    const vx_local_l1 *l1 = (const vx_local_l1 *)(h + 1);
    const uint64_t *leaves = (const uint64_t *)(l1 + h->n_l1);
    const uint32_t *attrs = (const uint32_t *)(leaves + h->n_leaves);
    const uint32_t *pal = attrs + h->n_attrs;
    vx_gpu_chunk &c = table[sl.chunk];
    // the old record's root, to carry the water lines of untouched L1 cells over
    const uint64_t old_mask = mk64(c.root[0], c.root[1]);
    const bool old_full = (c.root[2] & FULL_BIT) != 0u;
    const uint32_t old_run = c.root[2] & PTR_MASK;
    bool all_full = h->n_l1 == 64u;
    for (uint32_t i = 0; all_full && i < h->n_l1; i++) all_full = l1[i].full != 0u;
    const bool uniform = (h->flags & VX_LT_UNIFORM) || h->n_attrs == 0u;
    const uint32_t node_w = all_full ? 0u : 4u * h->n_l1;
    const uint32_t leaf_w0 = sl.offset + node_w;                 // leaf runs
    const uint32_t attr_w0 = leaf_w0 + 2u * h->n_leaves;          // attribute table
    const uint32_t pal_w0 = attr_w0 + (uniform ? 0u : h->n_attrs);  // palettes
    for (uint32_t i = lane; i < h->n_leaves; i += 64u) {
        flat[leaf_w0 + 2u * i] = (uint32_t)leaves[i];
        flat[leaf_w0 + 2u * i + 1u] = (uint32_t)(leaves[i] >> 32);
    }
    if (!uniform) {
        for (uint32_t i = lane; i < h->n_attrs; i += 64u) {
            const uint32_t e = attrs[i];
            flat[attr_w0 + i] = (e & UNIFORM_BIT) ? e : e + pal_w0;
        }
        for (uint32_t i = lane; i < h->n_palette; i += 64u) flat[pal_w0 + i] = pal[i];
    }
    uint32_t bit = 0;
    if (!all_full) {
        // the i-th node is L1 cell bit i-th set bit of root_mask (build order)
        for (uint32_t i = lane; i < h->n_l1; i += 64u) {
            uint64_t m = h->root_mask;
            for (uint32_t k = 0; k < i; k++) m &= m - 1ull;
            bit = (uint32_t)__ffsll((long long)m) - 1u;
            const bool full = l1[i].full != 0u;
            uint32_t wl;
            if ((sl.touched >> bit) & 1ull) {
                wl = full ? 15u : 0u;
            } else if (old_full) {
                wl = 15u;
            } else {
                const uint32_t oi = old_run + mask_below(old_mask, bit);
                wl = old_flat[4u * oi + 3u] >> 28;
            }
            uint32_t *o = flat + sl.offset + 4u * i;
            o[0] = (uint32_t)l1[i].mask;
            o[1] = (uint32_t)(l1[i].mask >> 32);
            o[2] = full ? FULL_BIT : (leaf_w0 / 2u + l1[i].leaf_start);
            o[3] = (l1[i].leaf_prefix & 0x0FFFFFFFu) | (wl << 28);
        }
    }
    if (lane != 0) return;
    if (h->n_l1 == 0u) {
        c.root[0] = c.root[1] = c.root[2] = c.root[3] = 0u;
    } else if (all_full) {
        c.root[0] = c.root[1] = 0xFFFFFFFFu;
        c.root[2] = FULL_BIT;
        c.root[3] = 0u;
    } else {
        c.root[0] = (uint32_t)h->root_mask;
        c.root[1] = (uint32_t)(h->root_mask >> 32);
        c.root[2] = sl.offset / 4u;
        c.root[3] = 0u;
    }
    uint32_t flags = (c.attr_flags & ~(ATTR_UNIFORM | ATTR_HAS_EMITTER)) | sl.flags;
    if (uniform) {
        flags |= ATTR_UNIFORM;
        c.attr_base = (h->flags & VX_LT_UNIFORM) ? h->uniform : 0u;
    } else {
        c.attr_base = attr_w0;
    }
    if (h->emitters) flags |= ATTR_HAS_EMITTER;
    c.attr_flags = flags;
    c.dry_mask[0] |= (uint32_t)sl.dry;
    c.dry_mask[1] |= (uint32_t)(sl.dry >> 32);
    if (sl.bmax[0] > 0) {  // voxels became solid: the box grows to them (set from them on an empty chunk)
        const bool was_empty = solid[sl.chunk] == 0u;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float lo = c.origin[k] + (float)sl.bmin[k], hi = c.origin[k] + (float)sl.bmax[k];
            c.aabb_min[k] = was_empty ? lo : fminf(c.aabb_min[k], lo);
            c.aabb_max[k] = was_empty ? hi : fmaxf(c.aabb_max[k], hi);
        }
    }
    solid[sl.chunk] = h->solid_count;
}

}  // namespace vxd

using namespace vxd;
using vxh::DevBuf;
using vxh::ensure;
using vxh::release;

struct vx_dworld {
    vx_ctx *ctx = nullptr;
    DevBuf table, solid;          // chunk records and solid voxel counts
    uint32_t n = 0;
    std::unordered_map<uint64_t, uint32_t> index;  // chunk key -> table index
    size_t used_words = 0, cap_words = 0;
    size_t live_words = 0;             // the used words right after the last compaction (or creation)
    DevBuf batch, dense, recs, slots;  // batch inputs and state, per-pass scratch, extract's slot
    uint8_t *h_stage = nullptr;        // pinned staging of the batch inputs and state
    size_t h_stage_cap = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    vx_dworld_stats st{};
    std::vector<uint32_t> h_solid;
};

namespace {

uint64_t key_of(int32_t cx, int32_t cy, int32_t cz) {
    return ((uint64_t)(uint32_t)cx << 42) ^ ((uint64_t)(uint32_t)cy << 21) ^ (uint64_t)(uint32_t)cz;
}

int32_t floordiv64(int32_t v) { return v >= 0 ? v / 64 : -((-v + 63) / 64); }

constexpr size_t SLOT_BYTES = (VX_LOCAL_TREE_MAX + 7u) & ~(size_t)7u;

void publish(vx_dworld *w) {
    w->ctx->flat_len = w->used_words * 4;
    w->ctx->pool_gen++;  // descriptors and captured graphs hold the pool sizes
}

size_t align16(size_t v) { return (v + 15u) & ~(size_t)15u; }

void release_all(vx_dworld *w) {
    DevBuf *bufs[] = {&w->table, &w->solid, &w->batch, &w->dense, &w->recs, &w->slots};
    for (DevBuf *b : bufs) release(*b);
    if (w->h_stage) (void)hipHostFree(w->h_stage);
    if (w->e0) (void)hipEventDestroy(w->e0);
    if (w->e1) (void)hipEventDestroy(w->e1);
    w->h_stage = nullptr;
    w->e0 = w->e1 = nullptr;
}

}  // namespace

namespace {

struct DwEdit {
    uint32_t slot, di;  // slot, dense index
    uint16_t id;
};

// The passes of one rebuild: the chunks slot_chunk (with the edits es, sorted by slot and
// voxel) decoded from old_flat (the context's pools), rebuilt and placed into dst from
// `used` words on. One staging image of every pass's inputs (slots, voxel runs, words =
// slot << 18 | dense index, ids) after the batch state, one copy, then the passes run
// without the host; passes take at most DW_GROUP chunks (dense ids and records are
// per-chunk scratch). out: the batch state after the passes; ms: their GPU time.
int run_passes(vx_dworld *w, const std::vector<uint32_t> &slot_chunk, const std::vector<DwEdit> &es,
               const uint32_t *old_flat, uint32_t *dst, size_t used, DwState &out, float &ms) {
    vx_ctx *c = w->ctx;
    const uint32_t n_slots = (uint32_t)slot_chunk.size();
    struct Pass { uint32_t s0, ns, n_runs; size_t slots, runs, words, ids; };
    std::vector<Pass> passes;
    std::vector<uint8_t> stage(align16(sizeof(DwState)));
    DwState st0{used, w->cap_words, 0ull, 0u, 0u};
    std::memcpy(stage.data(), &st0, sizeof st0);
    size_t eb = 0;
    for (uint32_t s0 = 0; s0 < n_slots; s0 += DW_GROUP) {
        Pass ps{};
        ps.s0 = s0;
        ps.ns = std::min(DW_GROUP, n_slots - s0);
        size_t ee = eb;
        while (ee < es.size() && es[ee].slot < s0 + ps.ns) ee++;
        std::vector<uint32_t> runs;
        for (size_t i = eb; i < ee; i++)
            if (i == eb || es[i].slot != es[i - 1].slot || es[i].di != es[i - 1].di) runs.push_back((uint32_t)(i - eb));
        runs.push_back((uint32_t)(ee - eb));
        ps.n_runs = (uint32_t)runs.size() - 1u;
        ps.slots = stage.size();
        ps.runs = align16(ps.slots + (size_t)ps.ns * sizeof(DwSlot));
        ps.words = align16(ps.runs + runs.size() * 4);
        ps.ids = align16(ps.words + (ee - eb) * 4);
        stage.resize(align16(ps.ids + (ee - eb) * 2));
        DwSlot *sl = (DwSlot *)(stage.data() + ps.slots);
        for (uint32_t k = 0; k < ps.ns; k++) {
            sl[k] = DwSlot{};
            sl[k].chunk = slot_chunk[s0 + k];
            for (int a = 0; a < 3; a++) { sl[k].bmin[a] = 64; sl[k].bmax[a] = 0; }
        }
        std::memcpy(stage.data() + ps.runs, runs.data(), runs.size() * 4);
        uint32_t *wd = (uint32_t *)(stage.data() + ps.words);
        uint16_t *id = (uint16_t *)(stage.data() + ps.ids);
        for (size_t i = eb; i < ee; i++) {
            wd[i - eb] = ((es[i].slot - s0) << 18) | es[i].di;
            id[i - eb] = es[i].id;
        }
        passes.push_back(ps);
        eb = ee;
    }
    const uint32_t max_ns = std::min(DW_GROUP, n_slots);
    int rc;
    if ((rc = ensure(w->batch, stage.size())) || (rc = ensure(w->dense, (size_t)max_ns * DW_VOL * 2)) ||
        (rc = ensure(w->recs, (size_t)max_ns * SLOT_BYTES)))
        return rc;
    if (w->h_stage_cap < stage.size()) {
        if (w->h_stage) (void)hipHostFree(w->h_stage);
        w->h_stage = nullptr;
        w->h_stage_cap = 0;
        if (hipHostMalloc((void **)&w->h_stage, stage.size(), 0) != hipSuccess) {
            w->h_stage = nullptr;
            return VX_E_OOM;
        }
        w->h_stage_cap = stage.size();
    }
    std::memcpy(w->h_stage, stage.data(), stage.size());
    uint8_t *d = (uint8_t *)w->batch.p;
    DwState *d_state = (DwState *)d;
    VX_HIPX(hipEventRecord(w->e0, c->stream));
    VX_HIPX(hipMemcpyAsync(d, w->h_stage, stage.size(), hipMemcpyHostToDevice, c->stream));
    Pools p = vxh::make_pools(c);
    p.chunks = (const vx_gpu_chunk *)w->table.p;
    p.n_chunks = w->n;
    for (const Pass &ps : passes) {
        DwSlot *slots = (DwSlot *)(d + ps.slots);
        k_dw_decode<<<ps.ns * 16u, 256, 0, c->stream>>>(p, slots, (uint16_t *)w->dense.p);
        if (ps.n_runs)
            k_dw_edits<<<(ps.n_runs + 255u) / 256u, 256, 0, c->stream>>>(
                (const uint32_t *)(d + ps.words), (const uint16_t *)(d + ps.ids), (const uint32_t *)(d + ps.runs),
                ps.n_runs, slots, (uint16_t *)w->dense.p);
        VX_LAUNCHED();
        if ((rc = vxh::build_local_async(c, (const uint16_t *)w->dense.p, ps.ns, (uint8_t *)w->recs.p, SLOT_BYTES)))
            break;
        k_dw_offsets<<<1, DW_GROUP, 0, c->stream>>>((const uint8_t *)w->recs.p, SLOT_BYTES, ps.ns, slots, d_state);
        k_dw_place<<<ps.ns, 64, 0, c->stream>>>((const uint8_t *)w->recs.p, SLOT_BYTES, slots, d_state, old_flat,
                                                dst, (vx_gpu_chunk *)w->table.p, (uint32_t *)w->solid.p);
        VX_LAUNCHED();
    }
    VX_HIPX(hipEventRecord(w->e1, c->stream));
    VX_HIPX(hipMemcpyAsync(w->h_stage, d_state, sizeof(DwState), hipMemcpyDeviceToHost, c->stream));
    VX_HIPX(hipMemcpyAsync(w->h_solid.data(), w->solid.p, (size_t)w->n * 4, hipMemcpyDeviceToHost, c->stream));
    VX_HIPX(hipStreamSynchronize(c->stream));
    if (rc) return rc;
    std::memcpy(&out, w->h_stage, sizeof out);
    VX_HIPX(hipEventElapsedTime(&ms, w->e0, w->e1));
    return VX_OK;
}

// Rebuild every non-empty chunk into a fresh buffer of the same capacity, packed from word
// 0 (the blocks of replaced and emptied chunks are dropped), then make it the pools.
// Records keep every field but their pointers: the rebuild replays no edit, so each L1
// cell keeps its water line and the box, dry mask and flags stay as they were.
int compact(vx_dworld *w) {
    vx_ctx *c = w->ctx;
    std::vector<uint32_t> chunks;
    for (uint32_t i = 0; i < w->n; i++)
        if (w->h_solid[i]) chunks.push_back(i);
    DevBuf nf;
    int rc;
    if ((rc = ensure(nf, w->cap_words * 4))) return rc;
    DwState st;
    float ms = 0.0f;
    rc = run_passes(w, chunks, {}, (const uint32_t *)c->flat.p, (uint32_t *)nf.p, 0, st, ms);
    if (rc == VX_OK && st.overflow) rc = VX_E_OVERFLOW;
    if (rc) {
        release(nf);
        return rc;
    }
    std::swap(c->flat, nf);
    release(nf);
    w->used_words = w->live_words = (size_t)st.used;
    w->st.compactions++;
    w->st.used_bytes = w->used_words * 4;
    publish(w);
    return VX_OK;
}

}  // namespace

extern "C" {

int vx_dworld_create(vx_ctx *c, const uint32_t *flat, size_t flat_words, size_t capacity_words,
                     const vx_gpu_chunk *table, const uint32_t *solid, uint32_t n, vx_dworld **out) {
    if (!c || !out || !flat || !table || !solid || n == 0 || flat_words % 32u || capacity_words < flat_words ||
        capacity_words * 4 > 0xFFFFFF00ull)
        return VX_E_ARG;
    for (uint32_t i = 0; i < n; i++)
        if (table[i].lod != 0u || table[i].voxel_size != 1.0f) return VX_E_ARG;  // LOD-0 chunks only
    VX_HIPX(hipSetDevice(c->device));
    vx_dworld *w = new (std::nothrow) vx_dworld();
    if (!w) return VX_E_OOM;
    w->ctx = c;
    int rc;
    if ((rc = ensure(c->flat, capacity_words * 4)) || (rc = ensure(w->table, (size_t)n * sizeof(vx_gpu_chunk))) ||
        (rc = ensure(w->solid, (size_t)n * 4))) {
        release_all(w);
        delete w;
        return rc;
    }
    if (hipEventCreate(&w->e0) != hipSuccess || hipEventCreate(&w->e1) != hipSuccess) {
        release_all(w);
        delete w;
        return VX_E_HIP;
    }
    VX_HIPX(hipMemcpyAsync(c->flat.p, flat, flat_words * 4, hipMemcpyHostToDevice, c->stream));
    VX_HIPX(hipMemcpyAsync(w->table.p, table, (size_t)n * sizeof(vx_gpu_chunk), hipMemcpyHostToDevice, c->stream));
    VX_HIPX(hipMemcpyAsync(w->solid.p, solid, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    VX_HIPX(hipStreamSynchronize(c->stream));
    w->n = n;
    w->h_solid.assign(solid, solid + n);
    for (uint32_t i = 0; i < n; i++)
        w->index[key_of((int32_t)floorf(table[i].origin[0] / 64.0f), (int32_t)floorf(table[i].origin[1] / 64.0f),
                        (int32_t)floorf(table[i].origin[2] / 64.0f))] = i;
    w->used_words = w->live_words = flat_words;
    w->cap_words = capacity_words;
    publish(w);
    *out = w;
    return VX_OK;
}

int vx_dworld_destroy(vx_dworld *w) {
    if (!w) return VX_E_ARG;
    (void)hipSetDevice(w->ctx->device);
    (void)hipStreamSynchronize(w->ctx->stream);
    release_all(w);
    w->ctx->flat_len = 0;  // the context reads its mirrors again
    w->ctx->pool_gen++;
    delete w;
    return VX_OK;
}

int vx_dworld_edit(vx_dworld *w, const vx_edit *edits, size_t n_edits) {
    if (!w || (n_edits && !edits)) return VX_E_ARG;
    vx_ctx *c = w->ctx;
    VX_HIPX(hipSetDevice(c->device));
    // Group the edits by chunk (slots in order of first appearance), then by voxel, keeping
    // each voxel's edits in their order; edits outside the world's chunks do nothing, as
    // set_block on a chunk that is not resident does nothing.
    std::vector<DwEdit> es;
    es.reserve(n_edits);
    std::vector<uint32_t> slot_chunk;
    std::unordered_map<uint32_t, uint32_t> slot_of;
    for (size_t i = 0; i < n_edits; i++) {
        const vx_edit &e = edits[i];
        const int32_t cx = floordiv64(e.x), cy = floordiv64(e.y), cz = floordiv64(e.z);
        auto it = w->index.find(key_of(cx, cy, cz));
        if (it == w->index.end()) continue;
        auto sit = slot_of.find(it->second);
        uint32_t s;
        if (sit == slot_of.end()) {
            s = (uint32_t)slot_chunk.size();
            slot_of.emplace(it->second, s);
            slot_chunk.push_back(it->second);
        } else {
            s = sit->second;
        }
        const uint32_t lx = (uint32_t)(e.x - cx * 64), ly = (uint32_t)(e.y - cy * 64), lz = (uint32_t)(e.z - cz * 64);
        es.push_back({s, lx | (lz << 6) | (ly << 12), e.id});
    }
    std::stable_sort(es.begin(), es.end(), [](const DwEdit &a, const DwEdit &b) {
        return a.slot != b.slot ? a.slot < b.slot : a.di < b.di;
    });
    const uint32_t n_slots = (uint32_t)slot_chunk.size();
    w->st.edits = n_edits;
    w->st.effective_edits = 0;
    w->st.chunks_rebuilt = 0;
    w->st.last_ms = 0.0f;
    if (n_slots == 0) return VX_OK;
    int rc;
    // Room for a batch that rebuilds the whole world half over: else compact first.
    if (w->used_words + w->live_words + w->live_words / 2 > w->cap_words && (rc = compact(w))) return rc;
    DwState st1;
    float ms = 0.0f;
    if ((rc = run_passes(w, slot_chunk, es, (const uint32_t *)c->flat.p, (uint32_t *)c->flat.p, w->used_words, st1,
                         ms)))
        return rc;
    w->st.last_ms = ms;
    w->st.effective_edits = st1.effective;
    w->st.chunks_rebuilt = st1.placed;  // a pass that did not fit stops the rest
    w->used_words = (size_t)st1.used;
    w->st.used_bytes = w->used_words * 4;
    w->st.capacity_bytes = w->cap_words * 4;
    publish(w);
    return st1.overflow ? VX_E_OVERFLOW : VX_OK;
}

int vx_dworld_compact(vx_dworld *w) {
    if (!w) return VX_E_ARG;
    VX_HIPX(hipSetDevice(w->ctx->device));
    return compact(w);
}

int vx_dworld_chunks(vx_dworld *w, vx_gpu_chunk *out, uint32_t cap, uint32_t *n_out) {
    if (!w || !n_out || (cap && !out)) return VX_E_ARG;
    vx_ctx *c = w->ctx;
    VX_HIPX(hipSetDevice(c->device));
    std::vector<vx_gpu_chunk> all(w->n);
    VX_HIPX(hipMemcpyAsync(all.data(), w->table.p, (size_t)w->n * sizeof(vx_gpu_chunk), hipMemcpyDeviceToHost,
                           c->stream));
    VX_HIPX(hipStreamSynchronize(c->stream));
    uint32_t k = 0;
    for (uint32_t i = 0; i < w->n; i++) {
        if (w->h_solid[i] == 0u) continue;  // empty chunks are not drawn (src/render/mod.rs:1463)
        if (k < cap) out[k] = all[i];
        k++;
    }
    *n_out = k;
    return k <= cap ? VX_OK : VX_E_OVERFLOW;
}

int vx_dworld_extract(vx_dworld *w, int32_t cx, int32_t cy, int32_t cz, uint16_t *dense) {
    if (!w || !dense) return VX_E_ARG;
    auto it = w->index.find(key_of(cx, cy, cz));
    if (it == w->index.end()) return VX_E_ARG;
    vx_ctx *c = w->ctx;
    VX_HIPX(hipSetDevice(c->device));
    int rc;
    if ((rc = ensure(w->slots, sizeof(DwSlot))) || (rc = ensure(w->dense, (size_t)DW_VOL * 2))) return rc;
    DwSlot s{};
    s.chunk = it->second;
    VX_HIPX(hipMemcpyAsync(w->slots.p, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
    Pools p = vxh::make_pools(c);
    p.chunks = (const vx_gpu_chunk *)w->table.p;
    p.n_chunks = w->n;
    k_dw_decode<<<16, 256, 0, c->stream>>>(p, (const DwSlot *)w->slots.p, (uint16_t *)w->dense.p);
    VX_LAUNCHED();
    VX_HIPX(hipMemcpyAsync(dense, w->dense.p, (size_t)DW_VOL * 2, hipMemcpyDeviceToHost, c->stream));
    VX_HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_dworld_stats_get(vx_dworld *w, vx_dworld_stats *out) {
    if (!w || !out) return VX_E_ARG;
    *out = w->st;
    out->used_bytes = w->used_words * 4;
    out->capacity_bytes = w->cap_words * 4;
    out->chunks = w->n;
    return VX_OK;
}

}  // extern "C"
