// lod.cpp -- LOD planning: which chunks of which LOD draw a camera's view, and the
// cross-fade weights of chunks changing LOD (C++ mirror of src/lod.rs:20-257).
//
// The plan is a quadtree walk over root columns of LOD max_lod around the camera: a
// chunk splits into its 8 children while its fine share (1 inside size * factor of its
// centre, 0 beyond that * fade_band, a smoothstep between) is positive. A split chunk
// whose children all draw is replaced by them; inside the fade band the children draw
// with fade f and the parent with fade -f, so that the per-pixel dither in the march
// (ray_skips) shows exactly one of the pair at every pixel.
#include <cmath>
#include <cstring>
#include <new>
#include <unordered_set>
#include <vector>

#include "vxhost.h"
#include "world.hpp"

using namespace vxh;

namespace {

struct Item {
    Key key;
    float fade;
    // RenderItem::share / is_fading (src/lod.rs:68-81)
    float share() const { return fade >= 0.0f ? fade : 1.0f + fade; }
    bool fading() const { return fade < 1.0f; }
};

using KeySet = std::unordered_set<Key, KeyHash>;

enum Residency { ABSENT = VXW_ABSENT, EMPTY = VXW_EMPTY, SOLID = VXW_SOLID };

struct Ctx {
    float cam[3];
    const vxw_lod_config *cfg;
    const World *world;
    vxw_residency_fn fn;
    void *user;
    const KeySet *was;
    KeySet *now;
    std::vector<Key> *desired;
    std::vector<Item> *render;

    Residency status(const Key &k) const {
        if (world) {
            auto it = world->chunks.find(k);
            if (it == world->chunks.end()) return ABSENT;
            return it->second.solid_count == 0 ? EMPTY : SOLID;
        }
        if (fn) {
            vxw_key q{{k.x, k.y, k.z}, (uint32_t)k.lod};
            int r = fn(user, &q);
            return r == VXW_SOLID ? SOLID : (r == VXW_EMPTY ? EMPTY : ABSENT);
        }
        return SOLID;
    }
};

// ChunkKey::bounds().distance_to(cam) (src/voxel/chunk.rs:28-31, src/math.rs:35-38):
// the length of the per-axis gap from the camera to the key's cube.
float bounds_distance(const Key &k, const float cam[3]) {
    const float size = (float)k.size();
    const float o[3] = {(float)(k.x * k.size()), (float)(k.y * k.size()), (float)(k.z * k.size())};
    float g[3];
    for (int i = 0; i < 3; i++) {
        const float lo = o[i] - cam[i], hi = cam[i] - (o[i] + size);
        g[i] = std::fmax(std::fmax(lo, hi), 0.0f);
    }
    return std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
}

// ChunkKey::center().distance(cam) (src/voxel/chunk.rs:32-34)
float centre_distance(const Key &k, const float cam[3]) {
    const float half = (float)k.size() * 0.5f;
    const float c[3] = {(float)(k.x * k.size()) + half, (float)(k.y * k.size()) + half,
                        (float)(k.z * k.size()) + half};
    const float d0 = c[0] - cam[0], d1 = c[1] - cam[1], d2 = c[2] - cam[2];
    return std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

// vacant (src/lod.rs:154-156)
bool vacant(const Key &k, const Ctx &c) {
    return k.y * k.size() >= WORLD_HEIGHT || bounds_distance(k, c.cam) > c.cfg->view_distance;
}

// fine_share (src/lod.rs:158-180)
float fine_share(const Key &k, const Ctx &c) {
    if (k.lod == 0) return 0.0f;
    const float inner = (float)k.size() * c.cfg->factor;
    const float outer = inner * c.cfg->fade_band;
    const float d = centre_distance(k, c.cam);
    if (!c.cfg->cross_fade) {
        const float threshold = c.was->count(k) ? outer : inner;
        return d < threshold ? 1.0f : 0.0f;
    }
    if (d <= inner) return 1.0f;
    if (d >= outer) return 0.0f;
    const float x = (outer - d) / (outer - inner);
    if (c.cfg->fade_schedule == VXW_FADE_SMOOTHERSTEP) return x * x * x * (x * (6.0f * x - 15.0f) + 10.0f);
    return x * x * (3.0f - 2.0f * x);
}

bool cover(const Key &k, Ctx &c, uint32_t depth);

// dissolve (src/lod.rs:221-232): a parent inside the fade band draws beside its children
// at the complementary fade, provided it is solid and none of the children already fades.
void dissolve(const Key &k, float f, size_t start, Ctx &c) {
    std::vector<Item> &r = *c.render;
    if (c.status(k) != SOLID || r.size() == start) return;
    for (size_t i = start; i < r.size(); i++)
        if (r[i].fading()) return;
    for (size_t i = start; i < r.size(); i++) r[i].fade = f;
    r.push_back(Item{k, -f});
}

// visit (src/lod.rs:182-219): true when the key's space is drawn by the plan.
bool visit(const Key &k, Ctx &c) {
    if (vacant(k, c)) return true;
    const float f = fine_share(k, c);
    if (f > 0.0f) {
        c.now->insert(k);
        if (f < 1.0f) c.desired->push_back(k);
        const size_t start = c.render->size();
        bool covered = true;
        for (int dz = 0; dz < 2; dz++)
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) covered &= visit(k.child(dx, dy, dz), c);
        if (covered) {
            if (f < 1.0f) dissolve(k, f, start, c);
            return true;
        }
        if (c.status(k) != ABSENT) {
            c.render->resize(start);
            c.render->push_back(Item{k, 1.0f});
            return true;
        }
        return false;
    }
    c.desired->push_back(k);
    return cover(k, c, c.cfg->cover_depth);
}

// cover (src/lod.rs:234-257): draw the key, or failing that its descendants down to
// depth levels, all or nothing.
bool cover(const Key &k, Ctx &c, uint32_t depth) {
    if (vacant(k, c)) return true;
    if (c.status(k) != ABSENT) {
        c.render->push_back(Item{k, 1.0f});
        return true;
    }
    if (k.lod == 0 || depth == 0) return false;
    const size_t start = c.render->size();
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++)
                if (!cover(k.child(dx, dy, dz), c, depth - 1)) {
                    c.render->resize(start);
                    return false;
                }
    return true;
}

}  // namespace

struct vxw_planner {
    KeySet split, scratch;
    std::vector<Key> desired;
    std::vector<Item> render;
};

extern "C" {

void vxw_lod_config_default(vxw_lod_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    // LodConfig::default (src/lod.rs:39-53)
    c->max_lod = 4;
    c->factor = 2.0f;
    c->view_distance = 2600.0f;
    c->fade_band = 1.25f;
    c->fade_schedule = VXW_FADE_SMOOTHSTEP;
    c->cover_depth = 2;
    c->cross_fade = 1;
    c->shadow_share = 1;
    c->offscreen_shadows = 1;
}

vxw_planner *vxw_planner_new(void) { return new (std::nothrow) vxw_planner(); }
void vxw_planner_free(vxw_planner *p) { delete p; }

// Planner::plan (src/lod.rs:108-152)
int vxw_planner_plan(vxw_planner *p, const float cam[3], const vxw_lod_config *cfg, const vxw_world *w,
                     vxw_residency_fn fn, void *user) {
    if (!p || !cam || !cfg || cfg->max_lod > VXW_MAX_LOD || !(cfg->view_distance >= 0.0f) ||
        cfg->fade_schedule > VXW_FADE_SMOOTHERSTEP)
        return VXW_E_ARG;
    p->desired.clear();
    p->render.clear();
    p->scratch.clear();
    const int32_t root = 64 << cfg->max_lod;
    const int32_t r = (int32_t)std::ceil(cfg->view_distance / (float)root) + 1;
    const int32_t cx = (int32_t)std::floor(cam[0] / (float)root);
    const int32_t cz = (int32_t)std::floor(cam[2] / (float)root);
    Ctx c{{cam[0], cam[1], cam[2]}, cfg, w ? &w->w : nullptr, fn, user, &p->split, &p->scratch, &p->desired,
          &p->render};
    for (int32_t dz = -r; dz <= r; dz++)
        for (int32_t dx = -r; dx <= r; dx++) visit(Key{cx + dx, 0, cz + dz, (int32_t)cfg->max_lod}, c);
    p->split.swap(p->scratch);
    return VXW_OK;
}

size_t vxw_planner_render(const vxw_planner *p, vxw_render_item *out, size_t cap) {
    if (!p) return 0;
    if (out)
        for (size_t i = 0; i < p->render.size() && i < cap; i++) {
            const Item &it = p->render[i];
            out[i].key = vxw_key{{it.key.x, it.key.y, it.key.z}, (uint32_t)it.key.lod};
            out[i].fade = it.fade;
        }
    return p->render.size();
}

size_t vxw_planner_desired(const vxw_planner *p, vxw_key *out, size_t cap) {
    if (!p) return 0;
    if (out)
        for (size_t i = 0; i < p->desired.size() && i < cap; i++) {
            const Key &k = p->desired[i];
            out[i] = vxw_key{{k.x, k.y, k.z}, (uint32_t)k.lod};
        }
    return p->desired.size();
}

}  // extern "C"
