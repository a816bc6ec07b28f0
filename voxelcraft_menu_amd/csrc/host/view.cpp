// view.cpp -- per-view setup: camera basis, frustum partition, chunk list, grid,
// and GpuFrame packing (C++ mirror of the reference host code listed per function).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "vxhost.h"
#include "world.hpp"

using namespace vxh;

namespace {

struct V3 {
    float x, y, z;
};
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
// glam Vec3::normalize: self * (1 / length).
inline V3 normalize(V3 a) { return mul(a, 1.0f / length(a)); }
inline V3 normalize_or(V3 a, V3 fb) {
    float r = 1.0f / length(a);
    if (std::isfinite(r) && r > 0.0f) return mul(a, r);
    return fb;
}
inline V3 vmax(V3 a, V3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }

struct Aabb {
    V3 min, max;
    // Aabb::distance_to (src/math.rs:35-38)
    float distance_to(V3 p) const { return length(vmax(vmax(sub(min, p), sub(p, max)), V3{0, 0, 0})); }
};

// Frustum::from_camera / contains_aabb (src/math.rs:69-117)
struct Frustum {
    V3 n[6];
    float d[6];
    Frustum(V3 pos, V3 fwd, V3 right, V3 up, float tan_half_fov, float aspect, float far_) {
        float hx = tan_half_fov * aspect, hy = tan_half_fov;
        auto mk = [&](int i, V3 v) {
            V3 nn = normalize(v);
            n[i] = nn;
            d[i] = -dot(nn, pos);
        };
        mk(0, add(mul(fwd, hx), right));
        mk(1, sub(mul(fwd, hx), right));
        mk(2, add(mul(fwd, hy), up));
        mk(3, sub(mul(fwd, hy), up));
        mk(4, fwd);
        V3 nf = mul(fwd, -1.0f);
        n[5] = nf;
        d[5] = -dot(nf, add(pos, mul(fwd, far_)));
    }
    bool contains(const Aabb &b) const {
        for (int i = 0; i < 6; i++) {
            V3 p{n[i].x > 0 ? b.max.x : b.min.x, n[i].y > 0 ? b.max.y : b.min.y, n[i].z > 0 ? b.max.z : b.min.z};
            if (dot(n[i], p) + d[i] < 0.0f) return false;
        }
        return true;
    }
};

// The fields of a ChunkRecord the chunk list reads (src/voxel/chunk.rs:73-111).
struct ViewRec {
    Key key;
    Inner root;
    uint32_t attr_base, attr_flags;  // GpuChunk attr_base / attr_flags (src/render/mod.rs:1485-1504)
    int32_t aabb_min[3], aabb_max[3];
    uint32_t solid_count;
    uint64_t dry_mask;
};

ViewRec view_rec(const ChunkRecord &r) {
    ViewRec v;
    v.key = r.key;
    v.root = r.root;
    v.attr_base = r.attr_base;
    uint32_t flags = r.attr_uniform ? 1u : 0u;
    if (r.has_water) flags |= 2u;
    if (r.has_foliage) flags |= 4u;
    if (r.has_cutout) flags |= 8u;
    if (r.has_emitter) flags |= 32u;
    if (r.has_glass) flags |= 16u;
    v.attr_flags = flags;
    for (int i = 0; i < 3; i++) {
        v.aabb_min[i] = r.aabb_min[i];
        v.aabb_max[i] = r.aabb_max[i];
    }
    v.solid_count = r.solid_count;
    v.dry_mask = r.dry_mask;
    return v;
}

Aabb world_aabb(const ViewRec &rec) {
    // ChunkRecord::world_aabb (src/voxel/chunk.rs:100-110): origin + local bounds * voxel size.
    const Key &k = rec.key;
    const float vs = (float)k.voxel_size();
    V3 o{(float)(k.x * k.size()), (float)(k.y * k.size()), (float)(k.z * k.size())};
    return Aabb{add(o, V3{(float)rec.aabb_min[0] * vs, (float)rec.aabb_min[1] * vs, (float)rec.aabb_min[2] * vs}),
                add(o, V3{(float)rec.aabb_max[0] * vs, (float)rec.aabb_max[1] * vs, (float)rec.aabb_max[2] * vs})};
}

// The 96-byte GpuChunk of a record (build_chunk_list, src/render/mod.rs:1506-1520).
void gpu_chunk_of(const ViewRec &rec, float fade, vx_gpu_chunk &g) {
    const Key &k = rec.key;
    const Aabb a = world_aabb(rec);
    std::memset(&g, 0, sizeof g);
    g.aabb_min[0] = a.min.x; g.aabb_min[1] = a.min.y; g.aabb_min[2] = a.min.z;
    g.voxel_size = (float)k.voxel_size();
    g.aabb_max[0] = a.max.x; g.aabb_max[1] = a.max.y; g.aabb_max[2] = a.max.z;
    g.lod = (uint32_t)k.lod;
    g.origin[0] = (float)(k.x * k.size()); g.origin[1] = (float)(k.y * k.size());
    g.origin[2] = (float)(k.z * k.size());
    g.fade = fade;
    g.root[0] = (uint32_t)rec.root.mask;
    g.root[1] = (uint32_t)(rec.root.mask >> 32);
    g.root[2] = rec.root.run_ptr;
    g.root[3] = rec.root.leaf_prefix;
    g.attr_base = rec.attr_base;
    g.attr_flags = rec.attr_flags;
    g.light_base = 0xF000u;  // LightRef::Uniform(0xF000): lighting is not on this path
    g.light_flags = 1u;
    g.dry_mask[0] = (uint32_t)rec.dry_mask;
    g.dry_mask[1] = (uint32_t)(rec.dry_mask >> 32);
}

// Frustum::from_camera for a view's camera (src/math.rs:69-117).
Frustum frustum_of(const vxw_camera *cam) {
    V3 pos{cam->pos[0], cam->pos[1], cam->pos[2]};
    V3 fwd{cam->fwd[0], cam->fwd[1], cam->fwd[2]};
    V3 right{cam->right[0], cam->right[1], cam->right[2]};
    V3 up{cam->up[0], cam->up[1], cam->up[2]};
    float tan_half = std::tan(cam->fov_y * 0.5f);
    float aspect = (float)cam->width / (float)cam->height;
    return Frustum(pos, fwd, right, up, tan_half, aspect, cam->far_);
}

struct DrawItem {
    const ViewRec *rec;
    float fade;
    float share() const { return fade >= 0.0f ? fade : 1.0f + fade; }
};

// f32::total_cmp's order as a signed integer.
int32_t total_order(float f) {
    int32_t b;
    std::memcpy(&b, &f, 4);
    return b ^ (int32_t)((uint32_t)(b >> 31) >> 1);
}

// partition_for_grid + build_chunk_list (src/render/mod.rs:82-138, :1447-1564).
// items: resident chunks with their fades, in draw-list order.
int build_view(const vxw_camera *cam, const std::vector<DrawItem> &items, bool shadow_share, bool offscreen_shadows,
               vx_gpu_chunk *chunks, uint32_t *grid, vxw_view *out) {
    V3 pos{cam->pos[0], cam->pos[1], cam->pos[2]};
    const Frustum fr = frustum_of(cam);
    // grid_origin (src/render/mod.rs:88-91)
    const int32_t gmin[3] = {(int32_t)std::floor(pos.x / 64.0f) - (int32_t)VX_GRID_X / 2, 0,
                             (int32_t)std::floor(pos.z / 64.0f) - (int32_t)VX_GRID_Z / 2};
    const int32_t gdim[3] = {(int32_t)VX_GRID_X, (int32_t)VX_GRID_Y, (int32_t)VX_GRID_Z};
    // grid_span (src/render/mod.rs:93-108): the grid cells a key covers, [lo, hi).
    auto grid_span = [&](const Key &k, int32_t lo[3], int32_t hi[3]) {
        const int32_t span = 1 << k.lod, p[3] = {k.x, k.y, k.z};
        bool any = true;
        for (int i = 0; i < 3; i++) {
            const int32_t base = p[i] * span - gmin[i];
            lo[i] = std::max(base, 0);
            hi[i] = std::min(base + span, gdim[i]);
            any = any && lo[i] < hi[i];
        }
        return any;
    };
    struct Item {
        float d;
        DrawItem it;
        const ViewRec *rec;
    };
    std::vector<Item> visible, offscreen;
    for (const DrawItem &it : items) {
        if (it.rec->solid_count == 0) continue;
        const ViewRec &rec = *it.rec;
        Aabb a = world_aabb(rec);
        int32_t lo[3], hi[3];
        if (fr.contains(a)) visible.push_back({a.distance_to(pos), it, &rec});
        else if (offscreen_shadows && grid_span(rec.key, lo, hi)) offscreen.push_back({a.distance_to(pos), it, &rec});
    }
    auto by_d = [](const Item &a, const Item &b) { return a.d < b.d; };
    std::stable_sort(visible.begin(), visible.end(), by_d);
    if (visible.size() > VX_MAX_CHUNKS) visible.resize(VX_MAX_CHUNKS);
    std::stable_sort(offscreen.begin(), offscreen.end(), by_d);
    if (offscreen.size() > VX_MAX_CHUNKS - visible.size()) offscreen.resize(VX_MAX_CHUNKS - visible.size());

    uint32_t camera_chunk = VX_NO_CHUNK;
    uint32_t idx = 0;
    std::vector<DrawItem> grid_set;
    grid_set.reserve(visible.size() + offscreen.size());
    auto emit = [&](const Item &it) {
        const ViewRec &rec = *it.rec;
        Aabb a = world_aabb(rec);
        if (idx < visible.size() && camera_chunk == VX_NO_CHUNK && pos.x >= a.min.x && pos.y >= a.min.y &&
            pos.z >= a.min.z && pos.x <= a.max.x && pos.y <= a.max.y && pos.z <= a.max.z)
            camera_chunk = idx;
        gpu_chunk_of(rec, it.it.fade, chunks[idx]);
        grid_set.push_back(it.it);
        idx++;
    };
    for (const auto &it : visible) emit(it);
    for (const auto &it : offscreen) emit(it);

    // Grid fill (mod.rs:1528-1555): a stable sort by (share, -lod), later items
    // overwriting earlier ones, so the largest share and then the finest LOD wins a cell.
    std::fill(grid, grid + VX_GRID_LEN, VX_NO_CHUNK);
    std::vector<uint32_t> order(grid_set.size());
    for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const float sa = shadow_share ? grid_set[a].share() : 1.0f;
        const float sb = shadow_share ? grid_set[b].share() : 1.0f;
        const int32_t ta = total_order(sa), tb = total_order(sb);
        if (ta != tb) return ta < tb;
        return grid_set[b].rec->key.lod < grid_set[a].rec->key.lod;
    });
    for (uint32_t i : order) {
        int32_t lo[3], hi[3];
        if (!grid_span(grid_set[i].rec->key, lo, hi)) continue;
        for (int32_t z = lo[2]; z < hi[2]; z++)
            for (int32_t y = lo[1]; y < hi[1]; y++)
                for (int32_t x = lo[0]; x < hi[0]; x++)
                    grid[(uint32_t)x + (uint32_t)y * VX_GRID_X + (uint32_t)z * VX_GRID_X * VX_GRID_Y] = i;
    }
    out->n_visible = (uint32_t)visible.size();
    out->n_total = idx;
    out->camera_chunk = camera_chunk;
    std::memcpy(out->grid_min, gmin, sizeof gmin);
    return VXW_OK;
}

}  // namespace

extern "C" void vxw_camera_basis(float yaw, float pitch, float fwd[3], float right[3], float up[3]) {
    float sy = std::sin(yaw), cy = std::cos(yaw), sp = std::sin(pitch), cp = std::cos(pitch);
    V3 f = normalize(V3{cy * cp, sp, sy * cp});
    V3 r = normalize_or(cross(f, V3{0, 1, 0}), V3{1, 0, 0});
    V3 u = normalize_or(cross(r, f), V3{0, 1, 0});
    fwd[0] = f.x; fwd[1] = f.y; fwd[2] = f.z;
    right[0] = r.x; right[1] = r.y; right[2] = r.z;
    up[0] = u.x; up[1] = u.y; up[2] = u.z;
}

extern "C" void vxw_sun(float t, float dir[3], float *daylight) {
    float a = (t - 0.25f) * 6.28318530717958647692f;
    V3 d = normalize(V3{std::cos(a) * 0.55f, std::sin(a), 0.42f});
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    if (daylight) *daylight = std::min(std::max((d.y + 0.12f) / 0.32f, 0.0f), 1.0f);
}

extern "C" uint32_t vxw_pairs_for(uint32_t tx, uint32_t ty) {
    uint64_t want = (uint64_t)tx * (uint64_t)ty * 16ull;
    want = std::max<uint64_t>(want, VX_MAX_PAIRS);
    want = std::min<uint64_t>(want, 0xFFFFFFFFull / 2);
    return (uint32_t)want;
}

namespace {

// The LOD-0 chunks in insertion order: the render set of vxw_build_view.
std::vector<ViewRec> lod0_records(const World &w) {
    std::vector<ViewRec> recs;
    recs.reserve(w.order.size());
    for (const Key &k : w.order)
        if (k.lod == 0) recs.push_back(view_rec(w.chunks.at(k)));
    return recs;
}

int build_view_all(const std::vector<ViewRec> &recs, const vxw_camera *cam, vx_gpu_chunk *chunks, uint32_t *grid,
                   vxw_view *out) {
    std::vector<DrawItem> items;
    items.reserve(recs.size());
    for (const ViewRec &r : recs) items.push_back(DrawItem{&r, 1.0f});
    return build_view(cam, items, true, true, chunks, grid, out);
}

}  // namespace

// build_chunk_list over the LOD-0 chunks in insertion order, fade 1.0, with
// offscreen_shadows and shadow_share on (src/lod.rs:40-52).
extern "C" int vxw_build_view(const vxw_world *wh, const vxw_camera *cam, vx_gpu_chunk *chunks, uint32_t *grid,
                              vxw_view *out) {
    if (!wh || !cam || !chunks || !grid || !out) return VXW_E_ARG;
    if (cam->width == 0 || cam->height == 0) return VXW_E_ARG;
    return build_view_all(lod0_records(wh->w), cam, chunks, grid, out);
}

extern "C" int vxw_build_view_items(const vxw_world *wh, const vxw_camera *cam, const vxw_render_item *items,
                                    size_t n, const vxw_lod_config *cfg, vx_gpu_chunk *chunks, uint32_t *grid,
                                    vxw_view *out) {
    if (!wh || !cam || !chunks || !grid || !out || (!items && n)) return VXW_E_ARG;
    if (cam->width == 0 || cam->height == 0) return VXW_E_ARG;
    vxw_lod_config def;
    if (!cfg) {
        vxw_lod_config_default(&def);
        cfg = &def;
    }
    std::vector<ViewRec> recs;
    std::vector<float> fades;
    recs.reserve(n);
    fades.reserve(n);
    for (size_t i = 0; i < n; i++) {
        const vxw_render_item &it = items[i];
        if (it.key.lod > VXW_MAX_LOD) return VXW_E_ARG;
        auto f = wh->w.chunks.find(Key{it.key.pos[0], it.key.pos[1], it.key.pos[2], (int32_t)it.key.lod});
        if (f == wh->w.chunks.end()) continue;  // not resident: build_chunk_list skips it
        recs.push_back(view_rec(f->second));
        fades.push_back(it.fade);
    }
    std::vector<DrawItem> v;
    v.reserve(recs.size());
    for (size_t i = 0; i < recs.size(); i++) v.push_back(DrawItem{&recs[i], fades[i]});
    return build_view(cam, v, cfg->shadow_share != 0, cfg->offscreen_shadows != 0, chunks, grid, out);
}

// A copy of the LOD-0 chunk records a view's chunk list reads, taken between edits, so
// that views can be set up while the world is being edited.
struct vxw_snapshot {
    std::vector<ViewRec> recs;
    uint32_t emitter_chunks = 0;
};

extern "C" vxw_snapshot *vxw_snapshot_new(void) { return new (std::nothrow) vxw_snapshot(); }
extern "C" void vxw_snapshot_free(vxw_snapshot *s) { delete s; }

extern "C" int vxw_snapshot_take(vxw_snapshot *s, const vxw_world *wh) {
    if (!s || !wh) return VXW_E_ARG;
    s->recs = lod0_records(wh->w);
    s->emitter_chunks = wh->w.emitter_chunks;
    return VXW_OK;
}

extern "C" uint32_t vxw_snapshot_emitter_chunks(const vxw_snapshot *s) { return s ? s->emitter_chunks : 0u; }

extern "C" size_t vxw_snapshot_chunks(const vxw_snapshot *s, vx_gpu_chunk *out, size_t cap) {
    if (!s) return 0;
    size_t n = 0;
    for (const ViewRec &r : s->recs) {
        if (r.solid_count == 0) continue;
        if (out && n < cap) gpu_chunk_of(r, 1.0f, out[n]);
        n++;
    }
    return n;
}

extern "C" size_t vxw_snapshot_chunks_all(const vxw_snapshot *s, vx_gpu_chunk *out, uint32_t *solid, size_t cap) {
    if (!s) return 0;
    size_t n = 0;
    for (const ViewRec &r : s->recs) {
        if (n < cap) {
            if (out) gpu_chunk_of(r, 1.0f, out[n]);
            if (solid) solid[n] = r.solid_count;
        }
        n++;
    }
    return n;
}

extern "C" int vxw_pack_camera_frame(const vxw_camera *cam, uint32_t max_pairs, vx_gpu_frame *f) {
    if (!cam || !f || cam->width == 0 || cam->height == 0) return VXW_E_ARG;
    vxw_view v{};
    v.n_visible = 0;
    v.n_total = 0;
    v.camera_chunk = VX_NO_CHUNK;
    // grid_origin (src/render/mod.rs:88-91)
    v.grid_min[0] = (int32_t)std::floor(cam->pos[0] / 64.0f) - (int32_t)VX_GRID_X / 2;
    v.grid_min[1] = 0;
    v.grid_min[2] = (int32_t)std::floor(cam->pos[2] / 64.0f) - (int32_t)VX_GRID_Z / 2;
    return vxw_pack_frame(cam, &v, max_pairs, f);
}

extern "C" int vxw_frustum_planes(const vxw_camera *cam, float planes[24]) {
    if (!cam || !planes || cam->width == 0 || cam->height == 0) return VXW_E_ARG;
    const Frustum fr = frustum_of(cam);
    for (int i = 0; i < 6; i++) {
        planes[4 * i] = fr.n[i].x;
        planes[4 * i + 1] = fr.n[i].y;
        planes[4 * i + 2] = fr.n[i].z;
        planes[4 * i + 3] = fr.d[i];
    }
    return VXW_OK;
}

extern "C" int vxw_build_view_snapshot(const vxw_snapshot *s, const vxw_camera *cam, vx_gpu_chunk *chunks,
                                       uint32_t *grid, vxw_view *out) {
    if (!s || !cam || !chunks || !grid || !out) return VXW_E_ARG;
    if (cam->width == 0 || cam->height == 0) return VXW_E_ARG;
    return build_view_all(s->recs, cam, chunks, grid, out);
}

// GpuFrame packing (src/render/mod.rs:1612-1710); shading-only fields stay zero.
extern "C" int vxw_pack_frame(const vxw_camera *cam, const vxw_view *view, uint32_t max_pairs, vx_gpu_frame *f) {
    if (!cam || !view || !f || cam->width == 0 || cam->height == 0) return VXW_E_ARG;
    std::memset(f, 0, sizeof *f);
    for (int i = 0; i < 3; i++) {
        f->cam_pos[i] = cam->pos[i];
        f->cam_fwd[i] = cam->fwd[i];
        f->cam_right[i] = cam->right[i];
        f->cam_up[i] = cam->up[i];
        f->sun_dir[i] = cam->sun_dir[i];
        f->grid_min[i] = view->grid_min[i];
    }
    f->time = cam->time;
    f->tan_half_fov = std::tan(cam->fov_y * 0.5f);
    f->aspect = (float)cam->width / (float)cam->height;
    f->far_ = cam->far_;
    f->daylight = cam->daylight;
    f->res[0] = cam->width;
    f->res[1] = cam->height;
    f->tiles[0] = (cam->width + 7) / 8;
    f->tiles[1] = (cam->height + 7) / 8;
    f->chunk_count = view->n_visible;
    f->camera_chunk = view->camera_chunk;
    f->flags = cam->flags;
    f->max_pairs = max_pairs ? max_pairs : vxw_pairs_for(f->tiles[0], f->tiles[1]);
    f->shadow_dist = 220.0f;
    f->sea_level = cam->sea_level;
    f->grid_dim[0] = VX_GRID_X;
    f->grid_dim[1] = VX_GRID_Y;
    f->grid_dim[2] = VX_GRID_Z;
    f->water_absorb = 1.0f;
    f->jitter[0] = cam->jitter[0];
    f->jitter[1] = cam->jitter[1];
    f->taa_feedback = 1.0f;
    f->frame_index = cam->frame_index % 64u;
    // the wave field's parameters, clamped as Renderer::render packs them (src/render/mod.rs:1690-1694)
    f->wave_amp = std::max(cam->waves[0], 0.0f);
    f->wave_scale = std::max(cam->waves[1], 1.0f);
    f->wave_speed = cam->waves[2];
    f->wave_reflect_slope = std::max(cam->waves[3], 0.0f);
    return VXW_OK;
}
