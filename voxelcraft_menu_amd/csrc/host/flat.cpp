// flat.cpp -- a LocalTree as one flat vx_local_tree record (the device build's output) and
// back: the host build to compare with, and the path from a device-built record into a
// world's pools.
#include <cstring>
#include <new>

#include "vxhost.h"
#include "world.hpp"

using namespace vxh;

namespace {

constexpr size_t HDR = sizeof(vx_local_tree), L1REC = sizeof(vx_local_l1);
static_assert(HDR == 64 && L1REC == 24, "record layout");

size_t record_bytes(size_t n_l1, size_t n_leaves, size_t n_attrs, size_t n_palette) {
    const size_t b = HDR + n_l1 * L1REC + n_leaves * 8 + n_attrs * 4 + n_palette * 4;
    return (b + 7) & ~(size_t)7;
}

}  // namespace

extern "C" {

int vxw_build_local_flat(const uint16_t *dense, void *out, size_t cap, size_t *written) {
    if (!dense || !written) return VXW_E_ARG;
    const LocalTree t = build_local(dense);
    const size_t bytes = record_bytes(t.l1.size(), t.leaves.size(), t.attrs.size(), t.palette.size());
    *written = bytes;
    if (!out) return VXW_OK;
    if (cap < bytes) return VXW_E_ARG;
    std::memset(out, 0, bytes);
    vx_local_tree h{};
    h.root_mask = t.root_mask;
    h.dry_mask = t.dry_mask;
    h.n_l1 = (uint32_t)t.l1.size();
    h.n_leaves = (uint32_t)t.leaves.size();
    h.n_attrs = (uint32_t)t.attrs.size();
    h.n_palette = (uint32_t)t.palette.size();
    h.uniform = t.has_uniform ? t.uniform : 0u;
    h.flags = (t.has_uniform ? VX_LT_UNIFORM : 0u) | (t.has_water ? VX_LT_WATER : 0u) |
              (t.has_foliage ? VX_LT_FOLIAGE : 0u) | (t.has_cutout ? VX_LT_CUTOUT : 0u) |
              (t.has_glass ? VX_LT_GLASS : 0u) | (t.has_emitter ? VX_LT_EMITTER : 0u);
    h.emitters = t.emitters;
    h.solid_count = t.solid_count;
    h.leaf_count = t.leaf_count;
    std::memcpy(h.aabb_min, t.aabb_min, 3);
    std::memcpy(h.aabb_max, t.aabb_max, 3);
    h.bytes = (uint32_t)bytes;
    uint8_t *o = (uint8_t *)out;
    std::memcpy(o, &h, HDR);
    o += HDR;
    for (const LocalL1 &n : t.l1) {
        vx_local_l1 r{n.mask, n.leaf_start, n.leaf_prefix, n.full ? 1u : 0u, n.water_line};
        std::memcpy(o, &r, L1REC);
        o += L1REC;
    }
    if (!t.leaves.empty()) std::memcpy(o, t.leaves.data(), t.leaves.size() * 8);
    o += t.leaves.size() * 8;
    if (!t.attrs.empty()) std::memcpy(o, t.attrs.data(), t.attrs.size() * 4);
    o += t.attrs.size() * 4;
    if (!t.palette.empty()) std::memcpy(o, t.palette.data(), t.palette.size() * 4);
    return VXW_OK;
}

int vxw_insert_local_flat(vxw_world *w, int32_t cx, int32_t cy, int32_t cz, const void *rec, size_t bytes) {
    if (!w || !rec || bytes < HDR) return VXW_E_ARG;
    vx_local_tree h;
    std::memcpy(&h, rec, HDR);
    if (h.n_l1 > 64 || h.n_leaves > 4096 || h.n_attrs > 4096 || h.n_palette > 4096u * 49u ||
        record_bytes(h.n_l1, h.n_leaves, h.n_attrs, h.n_palette) != h.bytes || h.bytes > bytes)
        return VXW_E_ARG;
    LocalTree t;
    t.root_mask = h.root_mask;
    t.dry_mask = h.dry_mask;
    t.has_uniform = (h.flags & VX_LT_UNIFORM) != 0;
    t.uniform = (BlockId)h.uniform;
    t.has_water = (h.flags & VX_LT_WATER) != 0;
    t.has_foliage = (h.flags & VX_LT_FOLIAGE) != 0;
    t.has_cutout = (h.flags & VX_LT_CUTOUT) != 0;
    t.has_glass = (h.flags & VX_LT_GLASS) != 0;
    t.has_emitter = (h.flags & VX_LT_EMITTER) != 0;
    t.emitters = h.emitters;
    t.solid_count = h.solid_count;
    t.leaf_count = h.leaf_count;
    std::memcpy(t.aabb_min, h.aabb_min, 3);
    std::memcpy(t.aabb_max, h.aabb_max, 3);
    const uint8_t *p = (const uint8_t *)rec + HDR;
    t.l1.resize(h.n_l1);
    for (uint32_t i = 0; i < h.n_l1; i++) {
        vx_local_l1 r;
        std::memcpy(&r, p, L1REC);
        p += L1REC;
        t.l1[i].mask = r.mask;
        t.l1[i].leaf_start = r.leaf_start;
        t.l1[i].leaf_prefix = r.leaf_prefix;
        t.l1[i].full = r.full != 0;
        t.l1[i].water_line = (uint8_t)r.water_line;
    }
    t.leaves.resize(h.n_leaves);
    if (h.n_leaves) std::memcpy(t.leaves.data(), p, (size_t)h.n_leaves * 8);
    p += (size_t)h.n_leaves * 8;
    t.attrs.resize(h.n_attrs);
    if (h.n_attrs) std::memcpy(t.attrs.data(), p, (size_t)h.n_attrs * 4);
    p += (size_t)h.n_attrs * 4;
    t.palette.resize(h.n_palette);
    if (h.n_palette) std::memcpy(t.palette.data(), p, (size_t)h.n_palette * 4);
    w->w.insert_local(Key{cx, cy, cz}, t);
    return VXW_OK;
}

}  // extern "C"
