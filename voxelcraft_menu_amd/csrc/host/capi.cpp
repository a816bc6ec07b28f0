// capi.cpp -- extern "C" surface of the host world store (declared in include/vxhost.h).
#include <cstring>
#include <new>

#include "vxhost.h"
#include "world.hpp"

using namespace vxh;

extern "C" {

vxw_world *vxw_world_new(void) { return new (std::nothrow) vxw_world(); }
void vxw_world_free(vxw_world *w) { delete w; }

int vxw_insert_dense(vxw_world *w, int32_t cx, int32_t cy, int32_t cz, const uint16_t *dense) {
    if (!w || !dense) return VXW_E_ARG;
    w->w.insert_local(Key{cx, cy, cz}, build_local(dense));
    return VXW_OK;
}

int vxw_remove_chunk(vxw_world *w, int32_t cx, int32_t cy, int32_t cz) {
    if (!w) return VXW_E_ARG;
    return w->w.remove_chunk(Key{cx, cy, cz}) ? VXW_OK : VXW_E_MISSING;
}

uint16_t vxw_get_block(const vxw_world *w, int32_t x, int32_t y, int32_t z) {
    return w ? w->w.get_block(x, y, z) : 0;
}

int vxw_set_block(vxw_world *w, int32_t x, int32_t y, int32_t z, uint16_t id) {
    if (!w) return VXW_E_ARG;
    return w->w.set_block(x, y, z, id) ? 1 : 0;
}

int64_t vxw_apply_edits(vxw_world *w, const vx_edit *edits, size_t n) {
    if (!w || (!edits && n)) return VXW_E_ARG;
    int64_t landed = 0;
    for (size_t i = 0; i < n; i++)
        if (w->w.set_block(edits[i].x, edits[i].y, edits[i].z, edits[i].id)) landed++;
    return landed;
}

int vxw_extract_dense(const vxw_world *w, int32_t cx, int32_t cy, int32_t cz, uint16_t *out) {
    if (!w || !out) return VXW_E_ARG;
    return w->w.extract_dense(Key{cx, cy, cz}, out) ? VXW_OK : VXW_E_MISSING;
}

static int chunk_info(const vxw_world *w, const Key &k, vxw_chunk_info *o) {
    if (!w || !o) return VXW_E_ARG;
    auto it = w->w.chunks.find(k);
    if (it == w->w.chunks.end()) return VXW_E_MISSING;
    const ChunkRecord &r = it->second;
    std::memset(o, 0, sizeof *o);
    o->key[0] = k.x; o->key[1] = k.y; o->key[2] = k.z;
    o->lod = (uint32_t)k.lod;
    o->root[0] = (uint32_t)r.root.mask;
    o->root[1] = (uint32_t)(r.root.mask >> 32);
    o->root[2] = r.root.run_ptr;
    o->root[3] = r.root.leaf_prefix;
    o->attr_uniform = r.attr_uniform;
    o->attr_base = r.attr_base;
    o->attr_len = r.attr_len;
    for (int i = 0; i < 3; i++) {
        o->aabb_min[i] = r.aabb_min[i];
        o->aabb_max[i] = r.aabb_max[i];
    }
    o->solid_count = r.solid_count;
    o->leaf_count = r.leaf_count;
    o->has_water = r.has_water;
    o->has_foliage = r.has_foliage;
    o->has_cutout = r.has_cutout;
    o->has_glass = r.has_glass;
    o->has_emitter = r.has_emitter;
    o->emitters = r.emitters;
    o->dry_mask = r.dry_mask;
    return VXW_OK;
}

int vxw_chunk_info_get(const vxw_world *w, int32_t cx, int32_t cy, int32_t cz, vxw_chunk_info *o) {
    return chunk_info(w, Key{cx, cy, cz}, o);
}

size_t vxw_chunk_count(const vxw_world *w) { return w ? w->w.chunks.size() : 0; }

size_t vxw_chunk_keys(const vxw_world *w, int32_t *keys, size_t cap) {
    if (!w) return 0;
    size_t n = 0;
    for (const Key &k : w->w.order) {
        if (k.lod != 0) continue;
        if (keys && n < cap) {
            keys[3 * n] = k.x;
            keys[3 * n + 1] = k.y;
            keys[3 * n + 2] = k.z;
        }
        n++;
    }
    return n;
}

// ---- chunks of any LOD ----

static bool key_ok(const vxw_key *k) { return k && k->lod <= VXW_MAX_LOD; }
static Key key_of(const vxw_key *k) { return Key{k->pos[0], k->pos[1], k->pos[2], (int32_t)k->lod}; }

int vxw_insert_dense_key(vxw_world *w, const vxw_key *k, const uint16_t *dense) {
    if (!w || !dense || !key_ok(k)) return VXW_E_ARG;
    w->w.insert_local(key_of(k), build_local(dense));
    return VXW_OK;
}

int vxw_remove_chunk_key(vxw_world *w, const vxw_key *k) {
    if (!w || !key_ok(k)) return VXW_E_ARG;
    return w->w.remove_chunk(key_of(k)) ? VXW_OK : VXW_E_MISSING;
}

int vxw_extract_dense_key(const vxw_world *w, const vxw_key *k, uint16_t *out) {
    if (!w || !out || !key_ok(k)) return VXW_E_ARG;
    return w->w.extract_dense(key_of(k), out) ? VXW_OK : VXW_E_MISSING;
}

int vxw_chunk_info_key(const vxw_world *w, const vxw_key *k, vxw_chunk_info *o) {
    if (!key_ok(k)) return VXW_E_ARG;
    return chunk_info(w, key_of(k), o);
}

size_t vxw_chunk_keys_all(const vxw_world *w, vxw_key *out, size_t cap) {
    if (!w) return 0;
    size_t n = w->w.order.size();
    if (out)
        for (size_t i = 0; i < n && i < cap; i++) {
            const Key &k = w->w.order[i];
            out[i] = vxw_key{{k.x, k.y, k.z}, (uint32_t)k.lod};
        }
    return n;
}

uint32_t vxw_emitter_chunks(const vxw_world *w) { return w ? w->w.emitter_chunks : 0; }

const void *vxw_pool_data(const vxw_world *w, int pool) {
    if (!w) return nullptr;
    switch (pool) {
        case VX_POOL_INNERS: return w->w.inners.data().data();
        case VX_POOL_LEAVES: return w->w.leaves.data().data();
        case VX_POOL_BRICKS: return w->w.bricks.data().data();
        default: return nullptr;
    }
}

size_t vxw_pool_bytes(const vxw_world *w, int pool) {
    if (!w) return 0;
    switch (pool) {
        case VX_POOL_INNERS: return w->w.inners.data().size() * 16;
        case VX_POOL_LEAVES: return w->w.leaves.data().size() * 8;
        case VX_POOL_BRICKS: return w->w.bricks.data().size() * 4;
        default: return 0;
    }
}

size_t vxw_take_dirty(vxw_world *w, int pool, vx_range *out, size_t cap) {
    if (!w) return 0;
    const std::vector<std::pair<uint32_t, uint32_t>> *d = nullptr;
    switch (pool) {
        case VX_POOL_INNERS: d = &w->w.inners.dirty(); break;
        case VX_POOL_LEAVES: d = &w->w.leaves.dirty(); break;
        case VX_POOL_BRICKS: d = &w->w.bricks.dirty(); break;
        default: return 0;
    }
    size_t n = d->size();
    if (!out) return n;
    std::vector<std::pair<uint32_t, uint32_t>> taken;
    switch (pool) {
        case VX_POOL_INNERS: taken = w->w.inners.take_dirty(); break;
        case VX_POOL_LEAVES: taken = w->w.leaves.take_dirty(); break;
        default: taken = w->w.bricks.take_dirty(); break;
    }
    size_t m = taken.size() < cap ? taken.size() : cap;
    for (size_t i = 0; i < m; i++) out[i] = vx_range{taken[i].first, taken[i].second};
    return m;
}

int vxw_pool_stats_get(const vxw_world *w, int pool, vxw_pool_stats *o) {
    if (!w || !o) return VXW_E_ARG;
    std::memset(o, 0, sizeof *o);
    switch (pool) {
        case VX_POOL_INNERS:
            o->live_elems = w->w.inners.live_elems;
            o->referenced_elems = w->w.inners.referenced_elems;
            o->live_runs = w->w.inners.live_runs();
            o->len = w->w.inners.data().size();
            return VXW_OK;
        case VX_POOL_LEAVES:
            o->live_elems = w->w.leaves.live_elems;
            o->referenced_elems = w->w.leaves.referenced_elems;
            o->live_runs = w->w.leaves.live_runs();
            o->len = w->w.leaves.data().size();
            return VXW_OK;
        case VX_POOL_BRICKS:
            o->live_elems = w->w.bricks.used_words;
            o->len = w->w.bricks.data().size();
            return VXW_OK;
        default: return VXW_E_ARG;
    }
}

}  // extern "C"
