// view.cpp -- per-view setup: camera basis, frustum partition, chunk list, grid,
// and GpuFrame packing (C++ mirror of the reference host code listed per function).
#include <algorithm>
#include <cmath>
#include <cstring>
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

Aabb world_aabb(const ChunkRecord &rec) {
    // ChunkRecord::world_aabb (src/voxel/chunk.rs:100-110), LOD 0: voxel size 1.
    V3 o{(float)(rec.key.x * 64), (float)(rec.key.y * 64), (float)(rec.key.z * 64)};
    return Aabb{add(o, V3{(float)rec.aabb_min[0], (float)rec.aabb_min[1], (float)rec.aabb_min[2]}),
                add(o, V3{(float)rec.aabb_max[0], (float)rec.aabb_max[1], (float)rec.aabb_max[2]})};
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

// partition_for_grid + build_chunk_list (src/render/mod.rs:82-138, :1447-1564), LOD 0,
// fade 1.0, offscreen_shadows and shadow_share on (src/lod.rs:40-52).
extern "C" int vxw_build_view(const vxw_world *wh, const vxw_camera *cam, vx_gpu_chunk *chunks, uint32_t *grid,
                              vxw_view *out) {
    if (!wh || !cam || !chunks || !grid || !out) return VXW_E_ARG;
    if (cam->width == 0 || cam->height == 0) return VXW_E_ARG;
    const World &w = wh->w;
    V3 pos{cam->pos[0], cam->pos[1], cam->pos[2]};
    V3 fwd{cam->fwd[0], cam->fwd[1], cam->fwd[2]};
    V3 right{cam->right[0], cam->right[1], cam->right[2]};
    V3 up{cam->up[0], cam->up[1], cam->up[2]};
    float tan_half = std::tan(cam->fov_y * 0.5f);
    float aspect = (float)cam->width / (float)cam->height;
    Frustum fr(pos, fwd, right, up, tan_half, aspect, cam->far_);
    // grid_origin (src/render/mod.rs:88-91)
    int32_t gmin[3] = {(int32_t)std::floor(pos.x / 64.0f) - (int32_t)VX_GRID_X / 2, 0,
                       (int32_t)std::floor(pos.z / 64.0f) - (int32_t)VX_GRID_Z / 2};
    auto grid_cell = [&](const Key &k, int32_t c[3]) {
        c[0] = k.x - gmin[0];
        c[1] = k.y - gmin[1];
        c[2] = k.z - gmin[2];
        return c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] < (int32_t)VX_GRID_X && c[1] < (int32_t)VX_GRID_Y &&
               c[2] < (int32_t)VX_GRID_Z;
    };
    struct Item {
        float d;
        const ChunkRecord *rec;
    };
    std::vector<Item> visible, offscreen;
    for (const Key &k : w.order) {
        const ChunkRecord &rec = w.chunks.at(k);
        if (rec.solid_count == 0) continue;
        Aabb a = world_aabb(rec);
        int32_t c[3];
        if (fr.contains(a)) visible.push_back({a.distance_to(pos), &rec});
        else if (grid_cell(k, c)) offscreen.push_back({a.distance_to(pos), &rec});
    }
    auto by_d = [](const Item &a, const Item &b) { return a.d < b.d; };
    std::stable_sort(visible.begin(), visible.end(), by_d);
    if (visible.size() > VX_MAX_CHUNKS) visible.resize(VX_MAX_CHUNKS);
    std::stable_sort(offscreen.begin(), offscreen.end(), by_d);
    if (offscreen.size() > VX_MAX_CHUNKS - visible.size()) offscreen.resize(VX_MAX_CHUNKS - visible.size());

    uint32_t camera_chunk = VX_NO_CHUNK;
    uint32_t idx = 0;
    std::fill(grid, grid + VX_GRID_LEN, VX_NO_CHUNK);
    auto emit = [&](const Item &it) {
        const ChunkRecord &rec = *it.rec;
        Aabb a = world_aabb(rec);
        if (idx < visible.size() && camera_chunk == VX_NO_CHUNK && pos.x >= a.min.x && pos.y >= a.min.y &&
            pos.z >= a.min.z && pos.x <= a.max.x && pos.y <= a.max.y && pos.z <= a.max.z)
            camera_chunk = idx;
        vx_gpu_chunk &g = chunks[idx];
        std::memset(&g, 0, sizeof g);
        g.aabb_min[0] = a.min.x; g.aabb_min[1] = a.min.y; g.aabb_min[2] = a.min.z;
        g.voxel_size = 1.0f;
        g.aabb_max[0] = a.max.x; g.aabb_max[1] = a.max.y; g.aabb_max[2] = a.max.z;
        g.lod = 0;
        g.origin[0] = (float)(rec.key.x * 64); g.origin[1] = (float)(rec.key.y * 64);
        g.origin[2] = (float)(rec.key.z * 64);
        g.fade = 1.0f;
        g.root[0] = (uint32_t)rec.root.mask;
        g.root[1] = (uint32_t)(rec.root.mask >> 32);
        g.root[2] = rec.root.run_ptr;
        g.root[3] = rec.root.leaf_prefix;
        uint32_t flags = rec.attr_uniform ? 1u : 0u;
        if (rec.has_water) flags |= 2u;
        if (rec.has_foliage) flags |= 4u;
        if (rec.has_cutout) flags |= 8u;
        if (rec.has_emitter) flags |= 32u;
        if (rec.has_glass) flags |= 16u;
        g.attr_base = rec.attr_base;
        g.attr_flags = flags;
        g.light_base = 0xF000u;  // LightRef::Uniform(0xF000): lighting is not on this path
        g.light_flags = 1u;
        g.dry_mask[0] = (uint32_t)rec.dry_mask;
        g.dry_mask[1] = (uint32_t)(rec.dry_mask >> 32);
        // Grid fill: every item has share 1.0 and LOD 0, so the sort by (share, -lod)
        // is the identity and a later item overwrites an earlier one (mod.rs:1531-1555).
        int32_t c[3];
        if (grid_cell(rec.key, c))
            grid[(uint32_t)c[0] + (uint32_t)c[1] * VX_GRID_X + (uint32_t)c[2] * VX_GRID_X * VX_GRID_Y] = idx;
        idx++;
    };
    for (const auto &it : visible) emit(it);
    for (const auto &it : offscreen) emit(it);
    out->n_visible = (uint32_t)visible.size();
    out->n_total = idx;
    out->camera_chunk = camera_chunk;
    std::memcpy(out->grid_min, gmin, sizeof gmin);
    return VXW_OK;
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
    return VXW_OK;
}
