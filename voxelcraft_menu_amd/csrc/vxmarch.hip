// vxmarch.hip -- extern "C" entry points of the ray-march library (include/vxmarch.h).
//
// A context owns one HIP stream and the device buffers the reference Renderer owns:
// the three tree pools (mirrored with GpuMirror::sync semantics), the chunk list and
// chunk grid, the visibility buffer, Hi-Z, the pair list, counters and the R16F
// shadow target.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "ctx.hpp"

using namespace vxd;
using vxh::DevBuf;
using vxh::ensure;

#define HIPX VX_HIPX
#define LAUNCHED VX_LAUNCHED

int vxh_hip_fail(hipError_t e) {
    if (e == hipErrorOutOfMemory) return VX_E_OOM;
    std::fprintf(stderr, "vxmarch: HIP error %d (%s)\n", (int)e, hipGetErrorString(e));
    return VX_E_HIP;
}
#define hip_fail vxh_hip_fail

namespace vxh {

int ensure(DevBuf &b, size_t bytes) {
    if (bytes <= b.cap) return VX_OK;
    if (b.p) HIPX(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    HIPX(hipMalloc(&b.p, bytes));
    b.cap = bytes;
    return VX_OK;
}

void release(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

int fill_u32(hipStream_t st, void *p, uint32_t v, size_t n) {
    if (!n) return VX_OK;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    k_fill_u32<<<blocks, 256, 0, st>>>((uint32_t *)p, v, n);
    HIPX(hipGetLastError());
    return VX_OK;
}

int fill_u64(hipStream_t st, void *p, uint64_t v, size_t n) {
    if (!n) return VX_OK;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    k_fill_u64<<<blocks, 256, 0, st>>>((uint64_t *)p, v, n);
    HIPX(hipGetLastError());
    return VX_OK;
}

void launch_order(hipStream_t st, const uint32_t *cost, uint32_t n, uint32_t *ctl, uint32_t *order) {
    unsigned nb = std::max(1u, std::min<unsigned>(ORDER_BLOCKS, (n + 255u) / 256u));
    k_cost_hist<<<nb, 256, 0, st>>>(cost, nullptr, n, nullptr, ctl);
    k_cost_scatter<<<nb, 256, 0, st>>>(cost, nullptr, n, nullptr, ctl, nullptr, nullptr, order);
}

}  // namespace vxh

static const size_t MIN_CAP = 1u << 20;  // GpuMirror MIN_CAP (src/render/pools.rs:11)

static uint32_t pairs_for(uint32_t tx, uint32_t ty) {
    uint64_t want = (uint64_t)tx * ty * 16ull;
    want = std::max<uint64_t>(want, VX_MAX_PAIRS);
    want = std::min<uint64_t>(want, 0xFFFFFFFFull / 2);
    return (uint32_t)want;
}

static int fill_u32(vx_ctx *c, void *p, uint32_t v, size_t n) { return vxh::fill_u32(c->stream, p, v, n); }
static int fill_u64(vx_ctx *c, void *p, uint64_t v, size_t n) { return vxh::fill_u64(c->stream, p, v, n); }

static int alloc_targets(vx_ctx *c, uint32_t w, uint32_t h) {
    if (w == 0 || h == 0) return VX_E_ARG;
    uint32_t tx = (w + 7) / 8, ty = (h + 7) / 8;
    if ((uint64_t)tx * ty > VX_MAX_TILES) return VX_E_OVERFLOW;  // warn_if_oversized
    // Outputs of the old size are gone: nothing reads them until the next frame.
    c->color_valid = c->image_valid = c->taa_valid = false;
    c->w = w;
    c->h = h;
    c->tx = tx;
    c->ty = ty;
    c->max_pairs = c->max_pairs_req ? c->max_pairs_req : pairs_for(tx, ty);
    int rc;
    if ((rc = ensure(c->vis, (size_t)w * h * 8)) || (rc = ensure(c->shadow, (size_t)w * h * 2)) ||
        (rc = ensure(c->hiz, (size_t)tx * ty * 4)))
        return rc;
    if (c->path == VX_PATH_PAIRS && (rc = ensure(c->pairs, (size_t)c->max_pairs * 2 * 4))) return rc;
    if ((rc = ensure(c->cost, (size_t)tx * ty * 4)) || (rc = ensure(c->order, (size_t)tx * ty * 4))) return rc;
    c->have_cost = false;
    if ((rc = fill_u64(c, c->vis.p, 0, (size_t)w * h)) || (rc = fill_u32(c, c->hiz.p, 0, (size_t)tx * ty))) return rc;
    if ((rc = fill_u32(c, c->shadow.p, 0, ((size_t)w * h + 1) / 2))) return rc;
    HIPX(hipStreamSynchronize(c->stream));
    c->have_frame = false;
    return VX_OK;
}

static unsigned pad8(unsigned n) { return (n + 7u) & ~7u; }

namespace vxh {

template <int O, int V>
static void launch_tiles_ov(bool hz, bool uw, unsigned nb, hipStream_t st, const TileArgs &a) {
    if (hz && !uw)
        k_march_tiles<O, V, true, false><<<nb, 64, 0, st>>>(a);
    else if (!hz && !uw)
        k_march_tiles<O, V, false, false><<<nb, 64, 0, st>>>(a);
    else if (hz)
        k_march_tiles<O, V, true, true><<<nb, 64, 0, st>>>(a);
    else
        k_march_tiles<O, V, false, true><<<nb, 64, 0, st>>>(a);
}

// Valid (occupancy, variant) pairs compiled into the library.
bool tuning_ok(uint32_t occ, uint32_t var) {
    if (occ == 4) return var == 0 || var == 1 || var == 4 || var == 5;
    if (occ == 6) return var == 0 || var == 4 || var == 5;
    return false;
}

// One workgroup (one wave) per tile, launched in `order`.
void launch_tiles(uint32_t occ, uint32_t var, bool hz, bool uw, unsigned nb, hipStream_t st, const TileArgs &a) {
    switch (occ * 16 + var) {
        case 4 * 16 + 1: launch_tiles_ov<4, 1>(hz, uw, nb, st, a); break;
        case 4 * 16 + 4: launch_tiles_ov<4, 4>(hz, uw, nb, st, a); break;
        case 4 * 16 + 5: launch_tiles_ov<4, 5>(hz, uw, nb, st, a); break;
        case 6 * 16 + 0: launch_tiles_ov<6, 0>(hz, uw, nb, st, a); break;
        case 6 * 16 + 4: launch_tiles_ov<6, 4>(hz, uw, nb, st, a); break;
        case 6 * 16 + 5: launch_tiles_ov<6, 5>(hz, uw, nb, st, a); break;
        default: launch_tiles_ov<4, 0>(hz, uw, nb, st, a); break;
    }
}

}  // namespace vxh

using vxh::TileArgs;
using vxh::launch_tiles;
using vxh::tuning_ok;

extern "C" {

const char *vx_error_string(int code) {
    switch (code) {
        case VX_OK: return "ok";
        case VX_E_NODEV: return "no HIP device";
        case VX_E_OOM: return "out of device memory";
        case VX_E_ARG: return "invalid argument";
        case VX_E_OVERFLOW: return "capacity exceeded";
        case VX_E_UNSUPPORTED: return "feature not supported by this library";
        case VX_E_HIP: return "HIP runtime error";
        case VX_E_BUSY: return "busy";
        default: return "unknown error";
    }
}

int vx_device_name(char *buf, size_t cap) {
    if (!buf || !cap) return VX_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return VX_E_NODEV;
    int dev = 0;
    HIPX(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPX(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, cap, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return VX_OK;
}

int vx_ctx_create(const vx_ctx_desc *d, vx_ctx **out) {
    if (!d || !out) return VX_E_ARG;
    *out = nullptr;
    if (d->path != VX_PATH_TILES && d->path != VX_PATH_PAIRS) return VX_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return VX_E_NODEV;
    vx_ctx *c = new (std::nothrow) vx_ctx();
    if (!c) return VX_E_OOM;
    int rc = VX_OK;
    do {
        if (d->device >= 0) {
            if (d->device >= n) { rc = VX_E_ARG; break; }
            if (hipSetDevice(d->device) != hipSuccess) { rc = VX_E_HIP; break; }
            c->device = d->device;
        } else if (hipGetDevice(&c->device) != hipSuccess) { rc = VX_E_HIP; break; }
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { rc = VX_E_HIP; break; }
        if ((rc = vxh::noalloc_setup(c->device))) break;
        for (auto &set : c->ring)
            for (auto &e : set)
                if (rc == VX_OK && hipEventCreate(&e) != hipSuccess) rc = VX_E_HIP;
        if (rc) break;
        c->ev = c->ring[0];
        c->path = d->path;
        c->max_pairs_req = d->max_pairs;
        if (const char *e = std::getenv("VX_TILES_OCC")) {
            int v = std::atoi(e);
            if (v == 4 || v == 6) c->occ = (uint32_t)v;
        }
        if (const char *e = std::getenv("VX_TILES_VARIANT")) {
            int v = std::atoi(e);
            if (v >= 0 && tuning_ok(c->occ, (uint32_t)v)) c->variant = (uint32_t)v;
        }
        for (auto &b : c->pool)
            if ((rc = ensure(b, MIN_CAP))) break;
        if (rc) break;
        if ((rc = ensure(c->grid, VX_GRID_LEN * 4)) || (rc = ensure(c->counters, CNT_TOTAL * 4)) ||
            (rc = ensure(c->chunks, 96 * 64)))
            break;
        if ((rc = fill_u32(c, c->grid.p, VX_NO_CHUNK, VX_GRID_LEN))) break;
        if ((rc = fill_u32(c, c->chunks.p, 0, 96 * 64 / 4))) break;
        if ((rc = alloc_targets(c, d->width, d->height))) break;
    } while (0);
    if (rc) {
        vx_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return VX_OK;
}

int vx_ctx_destroy(vx_ctx *c) {
    if (!c) return VX_E_ARG;
    if (c->n_batches > 0) return VX_E_BUSY;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf *bufs[] = {&c->pool[0], &c->pool[1],   &c->pool[2], &c->chunks,      &c->grid,     &c->vis,
                      &c->shadow,  &c->hiz,       &c->pairs,   &c->counters,    &c->cost,     &c->order,
                      &c->heat,    &c->build_dense, &c->build_leaf, &c->build_words, &c->build_l1, &c->build_out,
                      &c->atlas,   &c->faces,     &c->color,  &c->image,       &c->flat,
                      &c->hist[0], &c->hist[1],   &c->taa_f32,   &c->sec_r,       &c->sec_m,    &c->shaft,
                      &c->probe,   &c->soft_h,    &c->soft_mask};
    for (DevBuf *b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (auto &set : c->ring)
        for (auto &e : set)
            if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return VX_OK;
}

int vx_resize(vx_ctx *c, uint32_t w, uint32_t h) {
    if (!c) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    return alloc_targets(c, w, h);
}

// GpuMirror::sync (src/render/pools.rs:44-98): grow by doubling from 1 MiB with a
// full upload; otherwise merge dirty ranges with a 4096-byte gap and upload them,
// falling back to a full upload above 512 merged ranges.
int vx_pool_sync(vx_ctx *c, int pool, const void *data, size_t bytes, const vx_range *dirty, size_t n_dirty) {
    if (!c || pool < 0 || pool > 2 || (!data && bytes) || (!dirty && n_dirty)) return VX_E_ARG;
    static const size_t elem[3] = {16, 8, 4};
    if (bytes % elem[pool]) return VX_E_ARG;
    if (bytes > 0xFFFFFFF0ull) return VX_E_OVERFLOW;  // pools are read through 32-bit buffer ranges
    HIPX(hipSetDevice(c->device));
    DevBuf &b = c->pool[pool];
    const void *p0 = b.p;
    const size_t len0 = c->pool_len[pool];
    struct GenBump {
        vx_ctx *c;
        const DevBuf &b;
        const void *p0;
        size_t len0;
        int pool;
        ~GenBump() {
            if (b.p != p0 || c->pool_len[pool] != len0) c->pool_gen++;
        }
    } bump{c, b, p0, len0, pool};
    if (!c->pool_synced[pool] && bytes <= b.cap) {
        // A new context has no mirror contents yet: its first sync is a full upload
        // whatever the dirty list says (another mirror may have drained it already).
        if (bytes) HIPX(hipMemcpyAsync(b.p, data, bytes, hipMemcpyHostToDevice, c->stream));
        c->pool_uploaded[pool] += bytes;
        c->pool_len[pool] = bytes;
        c->pool_synced[pool] = true;
        HIPX(hipStreamSynchronize(c->stream));
        return VX_OK;
    }
    c->pool_synced[pool] = true;
    if (bytes > b.cap) {
        size_t cap = b.cap ? b.cap : MIN_CAP;
        while (cap < bytes) cap *= 2;
        HIPX(hipStreamSynchronize(c->stream));
        if (b.p) HIPX(hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
        HIPX(hipMalloc(&b.p, cap));
        b.cap = cap;
        HIPX(hipMemcpyAsync(b.p, data, bytes, hipMemcpyHostToDevice, c->stream));
        c->pool_uploaded[pool] += bytes;
        c->pool_len[pool] = bytes;
        HIPX(hipStreamSynchronize(c->stream));
        return VX_OK;
    }
    c->pool_len[pool] = bytes;
    if (n_dirty == 0) return VX_OK;
    std::vector<std::pair<uint64_t, uint64_t>> d(n_dirty);
    for (size_t i = 0; i < n_dirty; i++)
        d[i] = {(uint64_t)dirty[i].off * elem[pool], ((uint64_t)dirty[i].off + dirty[i].len) * elem[pool]};
    std::sort(d.begin(), d.end());
    std::vector<std::pair<uint64_t, uint64_t>> merged;
    for (auto &r : d) {
        if (!merged.empty() && r.first <= merged.back().second + 4096)
            merged.back().second = std::max(merged.back().second, r.second);
        else
            merged.push_back(r);
    }
    if (merged.size() > 512) {
        HIPX(hipMemcpyAsync(b.p, data, bytes, hipMemcpyHostToDevice, c->stream));
        c->pool_uploaded[pool] += bytes;
    } else {
        for (auto &r : merged) {
            uint64_t s = r.first & ~3ull, e = std::min<uint64_t>((r.second + 3) & ~3ull, bytes);
            if (e > s) {
                HIPX(hipMemcpyAsync((char *)b.p + s, (const char *)data + s, e - s, hipMemcpyHostToDevice,
                                    c->stream));
                c->pool_uploaded[pool] += e - s;
            }
        }
    }
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

// The per-chunk block layout (vxw_flat_layout): every kernel of this context reads its
// nodes, leaves and attribute words from these bytes until it is switched off (data NULL,
// bytes 0); the pool mirrors are kept and serve again from then on.
int vx_pool_set_flat(vx_ctx *c, const void *data, size_t bytes) {
    if (!c || (!data && bytes) || bytes % 128) return VX_E_ARG;
    if (bytes > 0xFFFFFFF0ull) return VX_E_OVERFLOW;
    HIPX(hipSetDevice(c->device));
    HIPX(hipStreamSynchronize(c->stream));
    if (bytes > c->flat.cap) {
        release(c->flat);
        if (int rc = ensure(c->flat, bytes)) return rc;
    }
    if (bytes) HIPX(hipMemcpyAsync(c->flat.p, data, bytes, hipMemcpyHostToDevice, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    c->flat_len = bytes;
    c->pool_gen++;  // descriptors and captured graphs hold pool addresses
    return VX_OK;
}

int vx_pool_mirror_get(vx_ctx *c, int pool, vx_pool_mirror *o) {
    if (!c || !o || pool < 0 || pool > 2) return VX_E_ARG;
    std::memset(o, 0, sizeof *o);
    o->cap = c->pool[pool].cap;
    o->len = c->pool_len[pool];
    o->uploaded_bytes = c->pool_uploaded[pool];
    o->generation = c->pool_gen;
    return VX_OK;
}

int vx_pool_read(vx_ctx *c, int pool, void *out, size_t bytes) {
    if (!c || pool < 0 || pool > 2 || (!out && bytes) || bytes > c->pool_len[pool]) return VX_E_ARG;
    if (!bytes) return VX_OK;
    HIPX(hipSetDevice(c->device));
    HIPX(hipMemcpyAsync(out, c->pool[pool].p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_set_chunks(vx_ctx *c, const vx_gpu_chunk *chunks, uint32_t n, const uint32_t *grid) {
    if (!c || (!chunks && n) || !grid || n > VX_MAX_CHUNKS) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    size_t bytes = (size_t)std::max<uint32_t>(n, 1) * sizeof(vx_gpu_chunk);
    if (bytes > c->chunks.cap) {
        HIPX(hipStreamSynchronize(c->stream));
        int rc = ensure(c->chunks, std::max<size_t>(bytes, c->chunks.cap * 2));
        if (rc) return rc;
    }
    if (n) HIPX(hipMemcpyAsync(c->chunks.p, chunks, (size_t)n * sizeof(vx_gpu_chunk), hipMemcpyHostToDevice, c->stream));
    HIPX(hipMemcpyAsync(c->grid.p, grid, VX_GRID_LEN * 4, hipMemcpyHostToDevice, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    c->n_chunks = n;
    return VX_OK;
}

}  // extern "C"

namespace vxh {

Pools make_pools(const vx_ctx *c) {
    Pools p;
    p.inners = (const uint4 *)c->pool[0].p;
    p.leaves = (const uint2 *)c->pool[1].p;
    p.bricks = (const uint32_t *)c->pool[2].p;
    p.chunks = (const vx_gpu_chunk *)c->chunks.p;
    p.grid = (const uint32_t *)c->grid.p;
    p.n_inners = (uint32_t)(c->pool_len[0] / 16);
    p.n_leaves = (uint32_t)(c->pool_len[1] / 8);
    p.n_bricks = (uint32_t)(c->pool_len[2] / 4);
    if (c->flat_len) {
        // One block per chunk: the three element kinds index the same buffer.
        p.inners = (const uint4 *)c->flat.p;
        p.leaves = (const uint2 *)c->flat.p;
        p.bricks = (const uint32_t *)c->flat.p;
        p.n_inners = (uint32_t)(c->flat_len / 16);
        p.n_leaves = (uint32_t)(c->flat_len / 8);
        p.n_bricks = (uint32_t)(c->flat_len / 4);
    }
    p.n_chunks = c->n_chunks;
    return p;
}

FrameK make_frame(const vx_gpu_frame *g) {
    FrameK f;
    std::memcpy(f.cam, g->cam_pos, 12);
    std::memcpy(f.fwd, g->cam_fwd, 12);
    std::memcpy(f.right, g->cam_right, 12);
    std::memcpy(f.up, g->cam_up, 12);
    std::memcpy(f.sun, g->sun_dir, 12);
    for (int i = 0; i < 3; i++) f.inv_sun[i] = 1.0f / g->sun_dir[i];  // IEEE division, as the kernels' 1.0 / rd
    f.tan_half_fov = g->tan_half_fov;
    f.aspect = g->aspect;
    f.far_ = g->far_;
    f.daylight = g->daylight;
    f.shadow_dist = g->shadow_dist;
    f.sea_level = g->sea_level;
    f.jitter[0] = g->jitter[0];
    f.jitter[1] = g->jitter[1];
    f.res[0] = g->res[0];
    f.res[1] = g->res[1];
    f.tiles[0] = g->tiles[0];
    f.tiles[1] = g->tiles[1];
    f.chunk_count = g->chunk_count;
    f.camera_chunk = g->camera_chunk;
    f.flags = g->flags;
    f.max_pairs = g->max_pairs;
    f.frame_index = g->frame_index;
    f.time = g->time;
    f.wave_amp = g->wave_amp;
    f.wave_scale = g->wave_scale;
    f.wave_speed = g->wave_speed;
    f.wave_reflect_slope = g->wave_reflect_slope;
    std::memcpy(f.grid_min, g->grid_min, 12);
    std::memcpy(f.grid_dim, g->grid_dim, 12);
    return f;
}

// Override values for a (flags & SPEC_MASK, spec_hi) pipeline key (src/render/mod.rs:2355-2689).
int make_spec(uint32_t spec, uint32_t spec_hi, uint32_t frame_flags, Spec &s) {
    if (spec & VX_FLAG_FOLIAGE) return VX_E_UNSUPPORTED;  // cross_quad samples the texture atlas
    s.bits = 0;
    if (spec & VX_FLAG_LEAF_CUTOUT) s.bits |= SPEC_CUTOUT;
    if (spec & VX_FLAG_GLASS) s.bits |= SPEC_GLASS;
    if (spec_hi & VX_FLAG_HI_FULL_MARCH) s.bits |= SPEC_FULL_MARCH;
    if (spec_hi & VX_FLAG_HI_EMITTER_WORLD) s.bits |= SPEC_EMITTER;
    if (!(spec_hi & VX_FLAG_HI_LEGACY_WATER)) s.bits |= SPEC_WATER_REPAIR;
    if (spec_hi & VX_FLAG_HI_GLASS_REFLECT) s.bits |= SPEC_GLASS_REFLECT;
    if (spec_hi & VX_FLAG_HI_SNELL_BEND) s.bits |= SPEC_SNELL_BEND;
    if (spec & VX_FLAG_WAVES) s.bits |= SPEC_WAVES;
    if (spec & VX_FLAG_WAVE_ANISO) s.bits |= SPEC_WAVE_ANISO;
    if (spec & VX_FLAG_WAVE_SHOAL) s.bits |= SPEC_WAVE_SHOAL;
    if (spec & VX_FLAG_WAVE_FILL) s.bits |= SPEC_WAVE_FILL;
    if (spec_hi & VX_FLAG_HI_WATER_LOOK) s.bits |= SPEC_WATER_LOOK;
    s.leaf_fill = (spec & VX_FLAG_LEAF_THIN) ? 0.62f : 0.72f;
    return VX_OK;
}

int check_frame(const vx_gpu_frame *g, uint32_t n_chunks) {
    if (g->res[0] == 0 || g->res[1] == 0) return VX_E_ARG;
    if (g->tiles[0] != (g->res[0] + 7) / 8 || g->tiles[1] != (g->res[1] + 7) / 8) return VX_E_ARG;
    if ((uint64_t)g->tiles[0] * g->tiles[1] > VX_MAX_TILES) return VX_E_OVERFLOW;
    if (g->chunk_count > n_chunks || g->chunk_count > VX_MAX_CHUNKS) return VX_E_ARG;
    if ((uint64_t)g->grid_dim[0] * g->grid_dim[1] * g->grid_dim[2] > VX_GRID_LEN) return VX_E_ARG;
    return VX_OK;
}

}  // namespace vxh

using vxh::make_frame;
using vxh::make_pools;
using vxh::make_spec;

extern "C" {

int vx_render(vx_ctx *c, const vx_gpu_frame *g, uint32_t spec, uint32_t spec_hi) {
    if (!c || !g) return VX_E_ARG;
    if (g->res[0] != c->w || g->res[1] != c->h) return VX_E_ARG;
    if (g->tiles[0] != c->tx || g->tiles[1] != c->ty) return VX_E_ARG;
    if (int rc0 = vxh::check_frame(g, c->n_chunks)) return rc0;
    // A FLAG_HEATMAP frame counts the iterations of every (tile, chunk) pair the reference
    // marches, so it runs the reference's pair decomposition whatever the context's path.
    const bool heatmap = (g->flags & VX_FLAG_HEATMAP) != 0u;
    const bool use_pairs = c->path == VX_PATH_PAIRS || heatmap;
    if (use_pairs && g->max_pairs != c->max_pairs) return VX_E_ARG;
    Spec s;
    int rc = make_spec(spec, spec_hi, g->flags, s);
    if (rc) return rc;
    HIPX(hipSetDevice(c->device));
    if (heatmap) {
        if ((rc = ensure(c->pairs, (size_t)c->max_pairs * 2 * 4)) || (rc = ensure(c->heat, (size_t)c->w * c->h * 4)))
            return rc;
        if ((rc = fill_u32(c, c->heat.p, 0, (size_t)c->w * c->h))) return rc;
    }
    FrameK f = make_frame(g);
    Pools p = make_pools(c);
    const bool hiz_on = (f.flags & VX_FLAG_HIZ) != 0u;
    const uint32_t ntiles = c->tx * c->ty;
    uint32_t *counters = (uint32_t *)c->counters.p;
    unsigned long long *vis = (unsigned long long *)c->vis.p;
    uint32_t *hiz = (uint32_t *)c->hiz.p;
    hipStream_t st = c->stream;

    c->ev = c->ring[c->ring_next];
    c->ring_next = (c->ring_next + 1) % vx_ctx::RING;
    if (c->ring_count < (uint32_t)vx_ctx::RING) c->ring_count++;
    HIPX(hipEventRecord(c->ev[0], st));
    if ((rc = fill_u32(c, counters, 0, CNT_TOTAL))) return rc;
    uint32_t *heat = heatmap ? (uint32_t *)c->heat.p : nullptr;
    if (use_pairs) {
        // clear vis (+ hiz when Hi-Z is off), tile_select, finalize, march,
        // [build_hiz, recover_select, finalize_recover, march] (src/render/mod.rs:1716-1795)
        if ((rc = fill_u64(c, vis, 0, (size_t)c->w * c->h))) return rc;
        if (!hiz_on && (rc = fill_u32(c, hiz, 0, ntiles))) return rc;
        uint32_t *pairs = (uint32_t *)c->pairs.p;
        dim3 sel_grid((c->tx + 15) / 16, (c->ty + 15) / 16);
        k_tile_select<<<sel_grid, 256, 0, st>>>(p, f, hiz, pairs, counters);
        LAUNCHED();
        k_finalize<<<1, 1, 0, st>>>(f, counters);
        LAUNCHED();
        HIPX(hipEventRecord(c->ev[1], st));
        k_march_pairs<<<2048, 256, 0, st>>>(p, f, s, pairs, counters, vis, heat);
        LAUNCHED();
        HIPX(hipEventRecord(c->ev[2], st));
        if (hiz_on) {
            k_build_hiz<<<(ntiles + 3) / 4, 256, 0, st>>>(f, vis, hiz);
            LAUNCHED();
            HIPX(hipEventRecord(c->ev[3], st));
            k_recover_select<<<(c->max_pairs + 255) / 256, 256, 0, st>>>(p, f, hiz, pairs, counters);
            LAUNCHED();
            k_finalize_recover<<<1, 1, 0, st>>>(f, counters);
            LAUNCHED();
            HIPX(hipEventRecord(c->ev[4], st));
            k_march_pairs<<<2048, 256, 0, st>>>(p, f, s, pairs, counters, vis, heat);
            LAUNCHED();
        } else {
            HIPX(hipEventRecord(c->ev[3], st));
            HIPX(hipEventRecord(c->ev[4], st));
        }
        HIPX(hipEventRecord(c->ev[5], st));
    } else {
        HIPX(hipEventRecord(c->ev[1], st));
        TileArgs a{p, f, s, hiz, vis, (uint16_t *)c->shadow.p, counters, (const uint32_t *)c->order.p,
                   (uint32_t *)c->cost.p};
        if (c->have_cost && c->ordered)
            vxh::launch_order(st, (const uint32_t *)c->cost.p, ntiles, counters + CNT_CTL, (uint32_t *)c->order.p);
        else
            k_iota_views<<<(ntiles + 255) / 256, 256, 0, st>>>((uint32_t *)c->order.p, nullptr, ntiles, nullptr, nullptr);
        LAUNCHED();
        launch_tiles(c->occ, c->variant, hiz_on, (f.flags & VX_FLAG_UNDERWATER) != 0u, ntiles, st, a);
        LAUNCHED();
        c->have_cost = true;
        HIPX(hipEventRecord(c->ev[2], st));
        HIPX(hipEventRecord(c->ev[3], st));
        HIPX(hipEventRecord(c->ev[4], st));
        HIPX(hipEventRecord(c->ev[5], st));
    }
    if (use_pairs || !(c->variant & VAR_FUSED_SHADOW)) {
        k_shadow<<<pad8((ntiles + 3) / 4), 256, 0, st>>>(p, f, s, vis, (uint16_t *)c->shadow.p, counters);
        LAUNCHED();
    }
    HIPX(hipEventRecord(c->ev[6], st));
    HIPX(hipGetLastError());
    c->frame = f;
    c->spec = s;
    c->gframe = *g;
    c->spec_word = spec;
    c->spec_hi_word = spec_hi;
    c->have_frame = true;
    c->heat_valid = heatmap;
    c->color_valid = false;
    return VX_OK;
}

int vx_sync(vx_ctx *c) {
    if (!c) return VX_E_ARG;
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

static int read_dev(vx_ctx *c, const DevBuf &b, void *out, size_t bytes) {
    HIPX(hipSetDevice(c->device));
    HIPX(hipMemcpyAsync(out, b.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_read_vis(vx_ctx *c, uint64_t *out, size_t n) {
    if (!c || !out || n != (size_t)c->w * c->h) return VX_E_ARG;
    return read_dev(c, c->vis, out, n * 8);
}

int vx_read_shadow(vx_ctx *c, uint16_t *out, size_t n) {
    if (!c || !out || n != (size_t)c->w * c->h) return VX_E_ARG;
    return read_dev(c, c->shadow, out, n * 2);
}

int vx_read_hiz(vx_ctx *c, uint32_t *out, size_t n) {
    if (!c || !out || n != (size_t)c->tx * c->ty) return VX_E_ARG;
    return read_dev(c, c->hiz, out, n * 4);
}

int vx_read_pairs(vx_ctx *c, uint32_t *out, size_t n, uint32_t *counters16) {
    if (!c || (!out && n) || !c->pairs.p || n > (size_t)c->max_pairs * 2) return VX_E_ARG;
    int rc = VX_OK;
    if (n && (rc = read_dev(c, c->pairs, out, n * 4))) return rc;
    if (counters16) rc = read_dev(c, c->counters, counters16, 16 * 4);
    return rc;
}

int vx_write_hiz(vx_ctx *c, const uint32_t *in, size_t n) {
    if (!c || !in || n != (size_t)c->tx * c->ty) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    HIPX(hipMemcpyAsync(c->hiz.p, in, n * 4, hipMemcpyHostToDevice, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_read_decoded(vx_ctx *c, vx_hit *out, size_t n) {
    if (!c || !out || n != (size_t)c->w * c->h) return VX_E_ARG;
    if (!c->have_frame) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    vx_hit *d = nullptr;
    HIPX(hipMallocAsync((void **)&d, n * sizeof(vx_hit), c->stream));
    k_decode<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(make_pools(c), c->frame,
                                                                 (const unsigned long long *)c->vis.p,
                                                                 (const uint16_t *)c->shadow.p, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, n * sizeof(vx_hit), hipMemcpyDeviceToHost, c->stream);
    (void)hipFreeAsync(d, c->stream);
    hipError_t e2 = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e);
    if (e2 != hipSuccess) return hip_fail(e2);
    return VX_OK;
}

int vx_pair_demand_get(vx_ctx *c, vx_pair_demand *out) {
    if (!c || !out) return VX_E_ARG;
    std::vector<uint32_t> cnt(CNT_LEN);
    int rc = read_dev(c, c->counters, cnt.data(), CNT_LEN * 4);
    if (rc) return rc;
    uint64_t st[4] = {0, 0, 0, 0};
    for (uint32_t sl = 0; sl < CNT_STAT_SLOTS; sl++)
        for (uint32_t k = 0; k < 4; k++) st[k] += cnt[CNT_STAT_BASE + sl * CNT_STAT_STRIDE + k];
    out->select = cnt[CNT_SELECT_RAW] + (uint32_t)st[STAT_SELECT];
    out->deferred = cnt[CNT_DEFERRED] + (uint32_t)st[STAT_DEFERRED];
    out->recover = cnt[CNT_RECOVER_RAW] + (uint32_t)st[STAT_RECOVER];
    out->cap = c->max_pairs;
    out->truncated = cnt[CNT_TRUNC];
    out->shadow_rays = (uint32_t)st[STAT_SHADOW];
    return VX_OK;
}

int vx_timings_get(vx_ctx *c, vx_timings *t) {
    if (!c || !t) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    HIPX(hipStreamSynchronize(c->stream));
    float ms[6];
    for (int i = 0; i < 6; i++) HIPX(hipEventElapsedTime(&ms[i], c->ev[i], c->ev[i + 1]));
    t->tile_select = ms[0];
    t->march = ms[1];
    t->hiz = ms[2];
    t->recover = ms[3];
    t->march2 = ms[4];
    t->shadow = ms[5];
    HIPX(hipEventElapsedTime(&t->total, c->ev[0], c->ev[6]));
    return VX_OK;
}

int vx_set_tuning(vx_ctx *c, const vx_tuning *t) {
    if (!c || !t) return VX_E_ARG;
    if (!tuning_ok(t->occupancy, t->variant)) return VX_E_ARG;
    if (t->ordered > 1) return VX_E_ARG;
    c->occ = t->occupancy;
    c->variant = t->variant;
    c->ordered = t->ordered;
    return VX_OK;
}

int vx_get_tuning(vx_ctx *c, vx_tuning *t) {
    if (!c || !t) return VX_E_ARG;
    t->occupancy = c->occ;
    t->variant = c->variant;
    t->ordered = c->ordered;
    return VX_OK;
}

int vx_timings_reset(vx_ctx *c) {
    if (!c) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    HIPX(hipStreamSynchronize(c->stream));
    c->ring_count = 0;
    return VX_OK;
}

int vx_timings_sum(vx_ctx *c, vx_timings *t, uint32_t *frames) {
    if (!c || !t) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    HIPX(hipStreamSynchronize(c->stream));
    std::memset(t, 0, sizeof *t);
    float *dst[6] = {&t->tile_select, &t->march, &t->hiz, &t->recover, &t->march2, &t->shadow};
    for (uint32_t k = 0; k < c->ring_count; k++) {
        uint32_t idx = (c->ring_next + vx_ctx::RING - 1 - k) % vx_ctx::RING;
        hipEvent_t *e = c->ring[idx];
        for (int i = 0; i < 6; i++) {
            float ms = 0;
            HIPX(hipEventElapsedTime(&ms, e[i], e[i + 1]));
            *dst[i] += ms;
        }
        float tot = 0;
        HIPX(hipEventElapsedTime(&tot, e[0], e[6]));
        t->total += tot;
    }
    if (frames) *frames = c->ring_count;
    return VX_OK;
}

int vx_read_heat(vx_ctx *c, uint32_t *out, size_t n) {
    if (!c || !out || n != (size_t)c->w * c->h || !c->heat_valid) return VX_E_ARG;
    return read_dev(c, c->heat, out, n * 4);
}

// Renderer::march_stats (src/render/mod.rs:1977-2045), on a host copy of the counter.
int vx_march_stats_compute(const uint32_t *heat, size_t n, uint32_t chunks, vx_march_stats *out) {
    if (!out || (!heat && n)) return VX_E_ARG;
    std::memset(out, 0, sizeof *out);
    std::vector<uint32_t> hot(heat, heat + n);
    uint64_t total = 0;
    for (uint32_t i : hot) {
        total += i;
        out->worst = std::max(out->worst, i);
        out->touched += i > 0u ? 1u : 0u;
        out->capped += i > MAX_ITERS ? 1u : 0u;
    }
    std::sort(hot.begin(), hot.end(), [](uint32_t a, uint32_t b) { return a > b; });
    auto share = [&](size_t k) -> double {
        k = std::min(k, hot.size());
        uint64_t s = 0;
        for (size_t i = 0; i < k; i++) s += hot[i];
        return total == 0 ? 0.0 : 100.0 * (double)s / (double)total;
    };
    auto pct = [&](double p) -> uint32_t {
        if (hot.empty()) return 0u;
        return hot[(size_t)(((100.0 - p) / 100.0) * (double)hot.size()) % hot.size()];
    };
    out->pixels = n;
    out->iters = total;
    out->top_1 = share(hot.size() / 100);
    out->top_10 = share(hot.size() / 10);
    out->p50 = pct(50.0);
    out->p90 = pct(90.0);
    out->p99 = pct(99.0);
    out->chunks = chunks;
    return VX_OK;
}

int vx_march_stats_get(vx_ctx *c, vx_march_stats *out) {
    if (!c || !out || !c->heat_valid) return VX_E_ARG;
    std::vector<uint32_t> heat((size_t)c->w * c->h);
    int rc = read_dev(c, c->heat, heat.data(), heat.size() * 4);
    if (rc) return rc;
    if ((rc = vx_march_stats_compute(heat.data(), heat.size(), c->frame.chunk_count, out))) return rc;
    return vx_pair_demand_get(c, &out->pairs);
}

int vx_device_outputs(vx_ctx *c, uint64_t **vis, uint16_t **shadow) {
    if (!c) return VX_E_ARG;
    if (vis) *vis = (uint64_t *)c->vis.p;
    if (shadow) *shadow = (uint16_t *)c->shadow.p;
    return VX_OK;
}

}  // extern "C"
