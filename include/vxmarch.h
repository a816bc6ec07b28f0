/*
 * vxmarch.h -- C ABI of the MI355X (gfx950) voxel ray-march library.
 *
 * Drop-in replacement for the GPU side of the reference renderer's hot path:
 * the primary-ray march writing the 64-bit visibility key, two-phase Hi-Z tile
 * culling, and the sun-shadow ray. Inputs are the reference's own byte layouts
 * (GpuFrame, GpuChunk, the three tree pools and the 32x8x32 chunk grid), so a
 * host that already builds them for wgpu hands the same bytes to these calls.
 *
 * Entry point                 replaces (reference file:line)
 *   vx_ctx_create/destroy     init_gpu + Renderer::new      (src/render/mod.rs:672-719, :851-1221)
 *   vx_resize                 Renderer::resize              (src/render/mod.rs:1232-1275)
 *   vx_pool_sync              GpuMirror::sync, via Renderer::sync_world
 *                                                           (src/render/pools.rs:44-98, src/render/mod.rs:1355-1396)
 *   vx_set_chunks             the chunk/grid write_buffer in build_chunk_list
 *                                                           (src/render/mod.rs:1555-1562)
 *   vx_render                 Renderer::render, passes tile_select..primary_shadow
 *                                                           (src/render/mod.rs:1593-1795, :1821-1840)
 *   vx_sync                   device.poll(Wait)             (src/headless.rs:489-496)
 *   vx_read_vis/shadow/hiz    read_texture / buffer readback (src/render/mod.rs:2062-2112)
 *   vx_read_decoded           shadow_guide + hit_t decode    (src/render/shaders/shadow.wgsl:8-30,
 *                                                            src/render/shaders/common.wgsl:1300-1318)
 *   vx_pair_demand            Renderer::pair_demand          (src/render/mod.rs:1948-1975)
 *   vx_march_stats / read_heat Renderer::march_stats           (src/render/mod.rs:1977-2045)
 *   vx_timings                Readback per-pass timestamps   (src/render/timer.rs:150-176)
 *   vx_resolve                the resolve pass, lane 0       (src/render/shaders/resolve.wgsl:1378-1466,
 *                                                            dispatched at src/render/mod.rs:1853-1878)
 *   vx_shaft_upload/scan      upload_shaft + the shaft scan  (src/render/mod.rs:1346-1352, :1726-1740)
 *   vx_probe_fill/write       make_probe + write_probe       (src/render/mod.rs:1398-1445, :2248-2315)
 *   vx_present                the present blit               (src/render/shaders/blit.wgsl:18-93)
 *   vx_batch_*                no counterpart: many views per launch (the reference renders
 *                             one view per render() call, src/headless.rs:680-837)
 *   vx_serve_*                no counterpart: continuous-batching request loop with
 *                             in-place edits (the reference's loop is headless.rs:419-497)
 *
 * Errors are int status codes; no exception crosses the ABI. A context owns one
 * HIP stream and is not re-entrant; several contexts may coexist.
 */
#ifndef VXMARCH_H
#define VXMARCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VX_OK = 0,
    VX_E_NODEV = -1,
    VX_E_OOM = -2,
    VX_E_ARG = -3,
    VX_E_OVERFLOW = -4,
    VX_E_UNSUPPORTED = -5,
    VX_E_HIP = -6,
    VX_E_BUSY = -8
};

/* Limits (src/render/mod.rs:20-37, :82-91). */
#define VX_TILE 8u
#define VX_MAX_CHUNKS 8191u
#define VX_MAX_PAIRS (1u << 20)
#define VX_MAX_TILES (1u << 19)
#define VX_GRID_X 32u
#define VX_GRID_Y 8u
#define VX_GRID_Z 32u
#define VX_GRID_LEN (VX_GRID_X * VX_GRID_Y * VX_GRID_Z)
#define VX_NO_CHUNK 0xFFFFFFFFu

/* Frame flag bits the passes read (src/render/mod.rs:144-270). */
#define VX_FLAG_HIZ 1u
#define VX_FLAG_HEATMAP 2u
#define VX_FLAG_SHADOWS 4u
#define VX_FLAG_UNDERWATER 64u
#define VX_FLAG_FOLIAGE 256u
#define VX_FLAG_LEAF_CUTOUT 65536u
#define VX_FLAG_GLASS 131072u
#define VX_FLAG_LEAF_THIN 524288u
#define VX_FLAG_WATER_FAR 2097152u
#define VX_FLAG_WATER_DARK 4194304u
#define VX_FLAG_WATER_REFLECT 16u
#define VX_FLAG_WATER_REFRACT 32u
#define VX_FLAG_WAVES 512u
#define VX_FLAG_WAVE_ANISO 1024u
#define VX_FLAG_WAVE_SHOAL 2048u
#define VX_FLAG_WAVE_FILL 4096u
#define VX_FLAG_WATER_SHADOW_CUT 262144u
#define VX_FLAG_DISTANT_SHADOWS 32768u
/* spec_hi bits (src/render/mod.rs:220-270). */
#define VX_FLAG_HI_FULL_MARCH 2u
#define VX_FLAG_HI_GLASS_REFLECT 128u
#define VX_FLAG_HI_WATER_LOOK 4096u
#define VX_FLAG_HI_EMITTER_WORLD 4u
#define VX_FLAG_HI_SNELL_BEND 256u
#define VX_FLAG_HI_WIND_SWAY 32768u
#define VX_FLAG_HI_LEGACY_WATER 2097152u

/* GpuFrame, 352 bytes (src/render/mod.rs:398-466). */
typedef struct vx_gpu_frame {
    float cam_pos[3];      float time;
    float cam_fwd[3];      float tan_half_fov;
    float cam_right[3];    float aspect;
    float cam_up[3];       float far_;
    float sun_dir[3];      float daylight;
    uint32_t res[2];       uint32_t tiles[2];
    uint32_t chunk_count;  uint32_t camera_chunk;
    uint32_t flags;        uint32_t max_pairs;
    float fog_density, fog_falloff, shadow_dist, ambient, fog_scatter, fog_g, fog_height;
    float mip_bias;
    int32_t grid_min[3];   float sea_level;
    uint32_t grid_dim[3];  float water_absorb;
    float jitter[2];       float taa_feedback;
    uint32_t frame_index;
    float cloud_cover, cloud_height, cloud_scale, cloud_speed, godray_strength;  /* offset 192 */
    uint32_t godray_steps; float godray_dist, cloud_shadow, tint_strength;
    int32_t tint_seed;     float tint_freq_t, tint_freq_h;
    float wave_amp, wave_scale, wave_speed, wave_reflect_slope;  /* offsets 240-255: the wave field */
    float shaft_origin[2], shaft_texel, shaft_soft, cloud_patch, cloud_relief;  /* offset 256 */
    float haze_warm, zenith_deep;
    float prev_view_proj[16];
} vx_gpu_frame;

/* GpuChunk, 96 bytes (src/render/mod.rs:468-489). */
typedef struct vx_gpu_chunk {
    float aabb_min[3];  float voxel_size;
    float aabb_max[3];  uint32_t lod;
    float origin[3];    float fade;
    uint32_t root[4];   /* Inner: mask lo, mask hi, run_ptr | FULL, leaf_prefix | water_line << 28 */
    uint32_t attr_base, attr_flags, light_base, light_flags;
    uint32_t dry_mask[2];
    uint32_t reserved[2];
} vx_gpu_chunk;

/* Dirty range in pool elements (RunPool/WordPool take_dirty, src/voxel/geometry.rs:213). */
typedef struct vx_range {
    uint32_t off, len;
} vx_range;

/* Edit journal record, 16 bytes (src/journal.rs:116-122). */
typedef struct vx_edit {
    int32_t x, y, z;
    uint16_t id, pad;
} vx_edit;

enum { VX_POOL_INNERS = 0, VX_POOL_LEAVES = 1, VX_POOL_BRICKS = 2 };

/* Algorithm used for the primary march.
 * VX_PATH_PAIRS: the reference's decomposition -- tile_select emits (tile, chunk)
 *   pairs, one wavefront marches one pair, then Hi-Z build/recover and a second march.
 * VX_PATH_TILES: one wavefront owns an 8x8 tile: culls the chunk list in registers,
 *   marches the surviving chunks front to back with exact per-lane early exit, and
 *   runs the same two-phase Hi-Z selection inside the wave. Same output bits. */
enum { VX_PATH_TILES = 0, VX_PATH_PAIRS = 1 };

typedef struct vx_ctx_desc {
    uint32_t width, height;   /* internal resolution */
    uint32_t max_pairs;       /* 0 = render::pairs_for(tiles) */
    uint32_t path;            /* VX_PATH_* */
    int32_t device;           /* HIP device ordinal, -1 = current */
} vx_ctx_desc;

typedef struct vx_ctx vx_ctx;

/* Counter readback (Renderer::pair_demand + march stats, src/render/mod.rs:1948-2045). */
typedef struct vx_pair_demand {
    uint32_t select, deferred, recover, cap;
    uint32_t truncated;       /* 1 if any list overflowed cap (TRUNCATION_MARK) */
    uint32_t shadow_rays;     /* pixels whose shadow_guide was valid */
} vx_pair_demand;

/* Per-pass GPU milliseconds of the last vx_render (hipEvent pairs on the context stream). */
typedef struct vx_timings {
    float tile_select, march, hiz, recover, march2, shadow, total;
} vx_timings;

/* Decoded visibility (voxel coordinate, face, material id, hit distance). */
typedef struct vx_hit {
    int32_t voxel[3];     /* world voxel (origin + v * voxel_size) */
    uint32_t face;        /* normal id 0..7, valid when t >= 0 */
    uint32_t material;    /* block id */
    uint32_t ci;          /* chunk index in the view list */
    float t;              /* hit distance (hit_t), -1 on a miss */
    uint32_t shadow;      /* 1 if the sun ray was blocked */
} vx_hit;

/* Kernel variants chosen by the autotuner (no counterpart in the reference). */
typedef struct vx_tuning {
    uint32_t occupancy;       /* tile march built for 4 or 6 waves per SIMD (6: variants 0, 4, 5) */
    uint32_t variant;         /* bit 0: classify the whole chunk list before marching;
                                 bit 2: trace the sun-shadow ray in the tile's wave (fused) */
    uint32_t ordered;         /* 1: launch tiles most expensive first (costs of the previous frame) */
} vx_tuning;

int vx_ctx_create(const vx_ctx_desc *desc, vx_ctx **out);
int vx_set_tuning(vx_ctx *ctx, const vx_tuning *t);
int vx_get_tuning(vx_ctx *ctx, vx_tuning *t);
int vx_ctx_destroy(vx_ctx *ctx);
int vx_resize(vx_ctx *ctx, uint32_t width, uint32_t height);
/* VX_E_OVERFLOW for a pool above 4 GiB - 16 bytes (kernels address pools with 32-bit
 * byte offsets). */
int vx_pool_sync(vx_ctx *ctx, int pool, const void *data, size_t bytes, const vx_range *dirty, size_t n_dirty);
/* A pool mirror's state (GpuMirror's cap and uploaded_bytes, src/render/pools.rs:1-9):
 * device capacity, bytes mirrored, bytes uploaded so far, and the context's pool
 * generation (bumped whenever a pool's device buffer or length changes). */
typedef struct vx_pool_mirror {
    uint64_t cap, len, uploaded_bytes;
    uint32_t generation, pad;
} vx_pool_mirror;
int vx_pool_mirror_get(vx_ctx *ctx, int pool, vx_pool_mirror *out);
/* Copies the first `bytes` of a pool's device mirror to the host (bytes <= its length);
 * no counterpart in the reference beyond a buffer readback. */
int vx_pool_read(vx_ctx *ctx, int pool, void *out, size_t bytes);
/* No counterpart in the reference: reads every pool from one buffer laid out as one
 * contiguous block per chunk (vxw_flat_layout in vxhost.h), for chunk records whose root
 * run pointer and attr_base that function rewrote. bytes is a multiple of 128; data NULL
 * with bytes 0 returns to the mirrors vx_pool_sync keeps. Batches and servers bound to
 * the context follow (their descriptors are rewritten before the next launch). */
int vx_pool_set_flat(vx_ctx *ctx, const void *data, size_t bytes);
int vx_set_chunks(vx_ctx *ctx, const vx_gpu_chunk *chunks, uint32_t n_total, const uint32_t *grid);
/* spec = frame flags & SPEC_MASK, spec_hi = spec_hi word (the pipeline key of
 * Renderer::render, src/render/mod.rs:1597). Foliage tufts need the texture atlas and
 * return VX_E_UNSUPPORTED. Under water with VX_FLAG_HI_SNELL_BEND, a ray that misses a
 * chunk going up is refracted at the sea surface and marched again through that chunk
 * (march.wgsl:23-45). */
int vx_render(vx_ctx *ctx, const vx_gpu_frame *frame, uint32_t spec, uint32_t spec_hi);
int vx_sync(vx_ctx *ctx);
int vx_read_vis(vx_ctx *ctx, uint64_t *out, size_t n);
int vx_read_shadow(vx_ctx *ctx, uint16_t *out, size_t n);
int vx_read_hiz(vx_ctx *ctx, uint32_t *out, size_t n);
int vx_write_hiz(vx_ctx *ctx, const uint32_t *in, size_t n);
/* Pair list + 16 counter words of the last frame (VX_PATH_PAIRS contexts only): the
 * reference's pairs/counters storage buffers (src/render/mod.rs:996-1037). */
int vx_read_pairs(vx_ctx *ctx, uint32_t *out, size_t n, uint32_t *counters16);
int vx_read_decoded(vx_ctx *ctx, vx_hit *out, size_t n);
int vx_pair_demand_get(vx_ctx *ctx, vx_pair_demand *out);
int vx_timings_get(vx_ctx *ctx, vx_timings *out);
/* Per-pass milliseconds summed over the frames rendered since vx_timings_reset (the
 * most recent 256 at most; *frames receives the count). */
int vx_timings_reset(vx_ctx *ctx);
int vx_timings_sum(vx_ctx *ctx, vx_timings *out, uint32_t *frames);
/* FLAG_HEATMAP (the reference's iteration heatmap, march.wgsl:51-53): a frame whose flags
 * hold VX_FLAG_HEATMAP runs the reference's pair decomposition and counts, per pixel, the
 * march_chunk iterations of every pair marched there (sun shadows are off in such frames,
 * as in shadow_active). vx_read_heat copies that counter (W*H u32); vx_march_stats_get
 * summarises it as Renderer::march_stats does (src/render/mod.rs:1977-2045). Both
 * return VX_E_ARG unless the last frame was a heatmap frame. */
typedef struct vx_march_stats {
    uint64_t pixels, iters, touched, capped;  /* capped: pixels above MAX_ITERS (512) */
    uint32_t worst, p50, p90, p99;
    double top_1, top_10;                     /* % of all iterations in the top 1 % / 10 % pixels */
    uint32_t chunks, pad;                     /* visible chunks of the frame */
    vx_pair_demand pairs;
} vx_march_stats;
int vx_read_heat(vx_ctx *ctx, uint32_t *out, size_t n);
int vx_march_stats_get(vx_ctx *ctx, vx_march_stats *out);
/* The same summary of a caller-held counter (no device work). */
int vx_march_stats_compute(const uint32_t *heat, size_t n, uint32_t chunks, vx_march_stats *out);

/* ---- Resolve, lanes 0, 1 and 2 ------------------------------------------------------
 * The reference's resolve passes (src/render/shaders/resolve.wgsl:1378-1466, the three
 * RESOLVE_LANE pipelines, with the water_sec pass before them, :1469-1535) over the last
 * vx_render frame, into an RGBA32F colour target (W*H*4 floats): a FLAG_HEATMAP
 * frame becomes heat(iterations / 300); a miss becomes the sky (sky_color without a cloud
 * deck); an opaque hit is shaded from its key as shade_hit_compact_cached /
 * shade_hit_legacy_cached do (resolve.wgsl:1537-2059: smooth light and AO from the 3x3
 * neighbourhood in front of the face, the sun visibility of the shadow passes -- 0 where
 * the shadow word holds a blocker distance, 1 elsewhere, shadow.wgsl:61-116 without soft
 * shadows --, sky specular, the atlas texel at the footprint's mip level) and mixed with
 * sky_base by the fog. A glass hit (lane 1) is shade_glass (:1261-1319): the transmit leg
 * traced and shaded (shade_water where it ends on water), the sky or, with
 * SPEC_GLASS_REFLECT, the traced reflect leg, mixed by schlick_glass. A water hit (lane 2)
 * is shade_water / shade_water_sec (:1078-1259): water_ctx (tint, medium light, wave
 * normals, gloss, roughness), the refract and reflect legs -- per pixel, or with
 * SPEC_WATER_SEC one pair per 2x2 block stored as f16 where water_share_safe holds --,
 * shore foam, glint and schlick. Leg hits are shaded by shade_hit_legacy with the sun
 * visibility of shadow_ray, without sky specular, with flat light under
 * SPEC_FLAT_SECONDARY. With FLAG_DISTANT_SHADOWS the sun visibility of a lit point is
 * multiplied by terrain_shade_beyond (common.wgsl:1571-1580) over the shaft field's last
 * vx_shaft_scan (zeros before any upload). The frame's shading fields (fog_*, ambient,
 * mip_bias, haze_warm, zenith_deep, water_absorb, the wave fields, shaft_origin,
 * shaft_texel, shaft_soft, tint_*) are read as given. With FLAG_TINT the albedo takes the
 * biome tint (resolve.wgsl:166-287: OpenSimplex2 temperature and humidity at the hit's
 * x, z, band weights, the tint table) weighted by the texel's alpha and tint_strength.
 * SPEC_CAUSTICS lifts the sun term of points up to 8 below the sea surface by the light
 * sheets of the first two wave octaves (resolve.wgsl:530-549). SPEC_SOFT_SHADOWS runs
 * shadow_horizontal and shadow_vertical (shadow.wgsl:49-116) over the frame's keys and
 * shadow words first -- per pixel a guided average of the occlusion over the row, then
 * the column, within the penumbra radius of each blocker distance (at most 4 pixels),
 * stored as an R8 mask -- and lane 0 takes its sun visibility from that mask.
 * FLAG_PROBE_TAP scales the gathered ambient by probe_tap (the probe field at fract(p / 512)
 * with clamp-to-edge linear filtering, resolve.wgsl:120-127); FLAG_PROBE_AMBIENT replaces it
 * by probe_tap and drops the face shade and the floor light.
 * Views that need what is not restated return VX_E_UNSUPPORTED: foliage-rich or
 * canopy-relief albedo, the water look, a cloud deck
 * (cloud_cover > 0), terrain god-ray shafts (godray_strength > 0 with
 * FLAG_TERRAIN_SHAFTS), the underwater view.
 * vx_resolve_set_assets copies the atlas and block face table the reference's host builds
 * (src/render/mod.rs:2159-2224, src/textures.rs; src/block.rs:96-102): RGBA8 sRGB texels,
 * mip level 0 of every layer, then level 1, ..., level l being (size >> l)^2 texels a
 * layer; 8 face words per block id (layer | perm class << 16). */
typedef struct vx_resolve_assets {
    const uint8_t *atlas;
    uint32_t size, layers, mips;
    const uint32_t *block_faces;
    uint32_t n_faces;             /* words */
} vx_resolve_assets;
int vx_resolve_set_assets(vx_ctx *ctx, const vx_resolve_assets *a);
int vx_resolve(vx_ctx *ctx);      /* asynchronous on the context stream */
/* Distant terrain shadows. vx_shaft_upload copies a 512 x 512 window of column heights
 * (Renderer::upload_shaft, src/render/mod.rs:1346-1352; n = 512 * 512, rows of z; the
 * host's vxw_shaft_* keeps the window). vx_shaft_scan runs shaft_scan's eight passes
 * (src/render/shaders/shaft.wgsl:1-24, encoded by Renderer::render when
 * FrameParams::shaft_scan is set, src/render/mod.rs:1726-1740) with the frame's sun_dir
 * and shaft_texel; vx_read_shaft copies the result (n = 512 * 512 floats). */
/* The ambient probe field (make_probe / write_probe, src/render/mod.rs:1398-1445,
 * :2240-2315): 128 x (6 x 128) x 128 RGBA16F texels -- x and z the probe lattice on a
 * 512-block torus, y the six cube-axis slabs of 128 levels -- texel (x, y, z) at
 * ((z * 768 + y) * 128 + x) * 4 halves, bounce rgb and occlusion in a. vx_probe_fill sets
 * every texel to *value (NULL: 1.0, the field a context starts with) or, with noise, to
 * make_probe's hashed values; a fill with a value pins the field (writes are then
 * ignored, as probe_fill pins it). vx_probe_write stores a chunk's bake (vxw_probe_bake:
 * occl 6 x 4096, bounce 6 x 4096 x 3) as f16 into its 16^3 block of each slab at
 * ((x / 4) mod 128, axis * 128 + clamp(y / 4, 0, 112), (z / 4) mod 128). vx_read_probe
 * copies the field (n = 128 * 768 * 128 * 4 halves). With FLAG_PROBE_CUBE or
 * FLAG_PROBE_BOUNCE the resolve samples it at hit + n * 2 (probe_field, trilinear,
 * repeating in x and z): the face shade takes its alpha, the floor light its rgb
 * (FLAG_PROBE_BOUNCE) and, with FLAG_PROBE_SUN, a sun bounce of its rgb. */
int vx_probe_fill(vx_ctx *ctx, const float *value, int noise);
int vx_probe_write(vx_ctx *ctx, const int32_t origin[3], const float *occl, const float *bounce);
int vx_read_probe(vx_ctx *ctx, uint16_t *out, size_t n);
int vx_shaft_upload(vx_ctx *ctx, const float *heights, size_t n);
int vx_shaft_scan(vx_ctx *ctx, const vx_gpu_frame *frame);
int vx_read_shaft(vx_ctx *ctx, float *out, size_t n);
int vx_read_color(vx_ctx *ctx, float *rgba, size_t n);  /* n = W*H*4 */
/* Present (src/render/shaders/blit.wgsl:18-93, src/render/blit.rs): the colour target
 * drawn onto a surface of the same size -- max(rgb, 0), the knee tonemap or the ACES fit
 * (tone 1), grade 1 or 2 mixed by strength below 1, the sRGB encode when the surface
 * format is not an sRGB one (encode 1) -- stored as RGBA8 (round(clamp(c) * 255), alpha
 * 255). vx_read_image copies it (n = W*H*4 bytes). */
typedef struct vx_present_desc {
    uint32_t encode, tone, grade;
    float strength;
} vx_present_desc;
int vx_present(vx_ctx *ctx, const vx_present_desc *d);
int vx_read_image(vx_ctx *ctx, uint8_t *rgba, size_t n);
/* TAA (src/render/shaders/taa.wgsl:1-127; the pass Renderer::render encodes after the
 * resolve when FrameParams::taa is set, src/render/mod.rs:1610-1620, :1884-1918): each
 * pixel's colour -- the resolve output as the reference's Rgba16Float trace target holds
 * it, rounded to f16 -- is mixed with the previous TAA output at the pixel's
 * reprojection through the frame's prev_view_proj (the key's hit point, or the view
 * direction for a miss), read with history_bicubic and clipped in YCoCg to the 3x3
 * neighbourhood's mean +- 1 sigma; the frame's taa_feedback is the weight of the colour
 * (1 keeps the colour, as the reference's first frame does). Heatmap frames keep the
 * colour. The output, f16 RGBA with alpha 1, is the next call's history, and vx_present
 * draws it until the next vx_resolve. vx_read_taa copies it (n = W*H*4 halves);
 * vx_write_taa_history replaces the history the next vx_taa reads. Histories start at
 * zero and are remade by a size change. */
int vx_taa(vx_ctx *ctx);
int vx_read_taa(vx_ctx *ctx, uint16_t *rgba, size_t n);
int vx_write_taa_history(vx_ctx *ctx, const uint16_t *rgba, size_t n);
/* Device pointers of the current outputs (for callers that keep data on the GPU). */
int vx_device_outputs(vx_ctx *ctx, uint64_t **vis, uint16_t **shadow);

/* ---- Rendering without allocation ---------------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device memory at
 * call time (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the tile march of vx_render (VX_PATH_TILES, tiles launched in raster
 * order) on device buffers the caller holds, on the calling thread's current HIP device
 * and the given stream (a hipStream_t, or NULL for the default stream). Inputs: the
 * frame (host memory, copied into the launch), the three pools, the chunk list and the
 * grid in the reference layouts; hiz (tiles words) is read as the previous frame's Hi-Z
 * and overwritten with this frame's; vis (W*H keys) and shadow (W*H R16F words) are
 * written whole. The output bits are vx_render's for the same inputs. The working memory
 * (counters, launch order, tile costs) is a fixed device array inside the library, one
 * per GPU, held until the process ends: at most VX_NOALLOC_MAX_TILES tiles per call
 * (VX_E_OVERFLOW beyond). The call allocates nothing and returns without waiting. Calls
 * that overlap on one GPU share that array and race: serialise them (one stream, or
 * stream order). FLAG_HEATMAP frames return VX_E_UNSUPPORTED. */
#define VX_NOALLOC_MAX_TILES (1u << 17)  /* 131072 tiles: 3840x2160 is 129600 */
typedef struct vx_device_view {
    const void *inners; uint64_t inners_bytes;  /* uint4 Inner nodes */
    const void *leaves; uint64_t leaves_bytes;  /* uint2 leaf masks */
    const void *bricks; uint64_t bricks_bytes;  /* u32 attribute words */
    const vx_gpu_chunk *chunks; uint32_t n_chunks;
    const uint32_t *grid;                       /* VX_GRID_LEN words */
} vx_device_view;
int vx_render_noalloc(const vx_gpu_frame *frame, uint32_t spec, uint32_t spec_hi, const vx_device_view *in,
                      uint32_t *hiz, uint64_t *vis, uint16_t *shadow, void *stream);

/* ---- View batches -----------------------------------------------------------------
 * No counterpart in the reference, which renders one view per Renderer::render call and
 * waits on it (src/headless.rs:680-837, :419-497). A batch renders many views of the
 * context's world in one submission. It shares the context's pools and stream (the
 * context must outlive it; vx_ctx_destroy returns VX_E_BUSY while batches exist), and
 * each view keeps its own Hi-Z, visibility and shadow buffers. Per view, the output bits
 * are those vx_render gives for the same frame, chunk list, grid and Hi-Z input. */
typedef struct vx_view {
    const vx_gpu_frame *frame;   /* the view's GpuFrame (res/tiles give its size) */
    const vx_gpu_chunk *chunks;  /* its chunk list, n_total records (build_chunk_list order) */
    uint32_t n_total;
    const uint32_t *grid;        /* its 32x8x32 chunk grid */
    uint32_t spec, spec_hi;      /* pipeline key, as for vx_render */
} vx_view;

/* PLAIN: one dispatch per view (the reference's pattern, kept as the baseline form);
 * FLAT: one launch, one workgroup per tile of every view, most expensive tiles first;
 * PERSISTENT: one launch of resident waves taking tiles from a global queue. */
enum { VX_BATCH_PERSISTENT = 0, VX_BATCH_FLAT = 1, VX_BATCH_PLAIN = 2 };

typedef struct vx_batch_tuning {
    uint32_t mode;            /* VX_BATCH_* */
    uint32_t occupancy;       /* waves per SIMD the march is built for: 4 or 6 */
    uint32_t variant;         /* vx_tuning.variant bits; batched forms need the fused shadow (4 or 5); they may add
                                 bit 3 (8): classify only the chunks on the candidate list of the tile's 64x64-pixel
                                 screen bin (built at the first render of the view set in such a form; a set
                                 whose lists would exceed 1 GiB, or VX_BIN_LIST_MAX_BYTES from the environment,
                                 keeps whole lists), not the view's whole list */
    uint32_t ordered;         /* 1: tiles most expensive first (their cost in the previous frame) */
    uint32_t queue_chunk;     /* PERSISTENT: tiles taken per queue grab, 1..64 */
    uint32_t waves_per_simd;  /* PERSISTENT: grid = CUs * 4 * waves_per_simd, 1..occupancy */
    uint32_t graph;           /* 1: capture the submission once into a hipGraph and replay it */
} vx_batch_tuning;

typedef struct vx_batch vx_batch;
typedef struct vx_serve vx_serve;

int vx_batch_create(vx_ctx *ctx, uint32_t max_views, vx_batch **out);
int vx_batch_destroy(vx_batch *b);
/* Replaces the view set (1..max_views views). Targets, Hi-Z and tile costs start at 0,
 * as for a new Renderer. */
int vx_batch_set_views(vx_batch *b, const vx_view *views, uint32_t n);

/* Chunk lists built on the device (no counterpart in the reference, which builds each
 * view's list on the CPU: partition_for_grid + build_chunk_list, src/render/mod.rs:82-138,
 * :1447-1564). vx_batch_set_chunk_table uploads the LOD-0 GpuChunk records of a render
 * set (fade 1, in render-set order, at most VX_CHUNK_TABLE_MAX; e.g. vxw_snapshot_chunks);
 * call it again after edits, before the next vx_batch_set_cameras. vx_batch_set_cameras
 * then sets views from cameras alone: each view's frame as vxw_pack_frame gives it (its
 * chunk_count and camera_chunk are ignored) and its frustum planes (vxw_frustum_planes);
 * the device culls, sorts and packs the view's chunk list and grid, which are the bytes
 * vxw_build_view gives for that camera. vx_batch_view_counts / vx_batch_read_chunks return
 * what was built. */
#define VX_CHUNK_TABLE_MAX 8192u
typedef struct vx_camera_view {
    const vx_gpu_frame *frame;
    uint32_t spec, spec_hi;
    float planes[24];           /* (nx, ny, nz, d) x 6 */
} vx_camera_view;
int vx_batch_set_chunk_table(vx_batch *b, const vx_gpu_chunk *chunks, uint32_t n);
int vx_batch_set_cameras(vx_batch *b, const vx_camera_view *views, uint32_t n);
/* The same views one frame later, their cameras moved: as vx_batch_set_cameras, but each
 * view keeps its Hi-Z words and tile costs, as the reference's Renderer keeps its Hi-Z
 * buffer while the camera moves from frame to frame (src/render/mod.rs:1716-1795, driven
 * by src/headless.rs:709-733). The views must be the installed ones in count, order, size
 * and class (Hi-Z on/off, camera under water or not); VX_E_ARG otherwise. */
int vx_batch_move_cameras(vx_batch *b, const vx_camera_view *views, uint32_t n);
int vx_batch_view_counts(vx_batch *b, uint32_t view, uint32_t *n_total, uint32_t *n_visible, uint32_t *camera_chunk);
int vx_batch_read_chunks(vx_batch *b, uint32_t view, vx_gpu_chunk *chunks, size_t cap, uint32_t *grid);
int vx_batch_set_tuning(vx_batch *b, const vx_batch_tuning *t);
int vx_batch_get_tuning(vx_batch *b, vx_batch_tuning *t);
/* Enqueues one frame of every view on the context stream (asynchronous). */
int vx_batch_render(vx_batch *b);
int vx_batch_sync(vx_batch *b);
/* Times each batched form for reps frames and keeps the fastest (the autotuner). */
int vx_batch_autotune(vx_batch *b, uint32_t reps, vx_batch_tuning *best, float *best_ms);
int vx_batch_reset_hiz(vx_batch *b);
int vx_batch_read_vis(vx_batch *b, uint32_t view, uint64_t *out, size_t n);
int vx_batch_read_shadow(vx_batch *b, uint32_t view, uint16_t *out, size_t n);
int vx_batch_read_hiz(vx_batch *b, uint32_t view, uint32_t *out, size_t n);
/* Totals over the views of the last frame (select/deferred/recover/shadow_rays). */
int vx_batch_demand_get(vx_batch *b, vx_pair_demand *out);
/* GPU milliseconds summed over the submissions since vx_batch_timings_reset (the last
 * 256 at most): the whole submission, and the march kernel alone (FLAT / PERSISTENT
 * without graph replay; otherwise the same span as the total). */
int vx_batch_timings_reset(vx_batch *b);
int vx_batch_timings_sum(vx_batch *b, float *total_ms, float *march_ms, uint32_t *frames);
int vx_batch_device_outputs(vx_batch *b, uint32_t view, uint64_t **vis, uint16_t **shadow);

/* Secondary bounce (the reflect / refract / transmit legs of resolve.wgsl's shade_water
 * and shade_glass, src/render/shaders/resolve.wgsl:1029-1058, :1209-1220, :1261-1310, each
 * traced with trace_world, common.wgsl:1205-1253). With it on, every vx_batch_render also
 * gathers the legs of the pixels whose primary hit is water or (glass specialisation)
 * glass into compact queues and traces them. The water reflect leg follows water_ctx's
 * mirror_t (resolve.wgsl:1078-1167): on top faces of views whose pipeline key and frame
 * flags enable waves (VX_FLAG_WAVES, with VX_FLAG_WAVE_ANISO / _SHOAL / _FILL) it reflects
 * about the wave normal at wave_reflect_slope, from the frame's time and wave_amp /
 * wave_scale / wave_speed; views with the water-look ripples (VX_FLAG_HI_WATER_LOOK)
 * return VX_E_UNSUPPORTED. Four records per pixel: [0] refract (water) or transmit
 * (glass), [1] reflect, and where the glass transmit leg ends on water, shade_water's legs
 * from that point (resolve.wgsl:1281-1289): [2] refract, [3] reflect. At each hit that
 * shade_hit would shade, the sun test it makes runs too (resolve.wgsl:741-745:
 * dot(n, sun) > 0, daylight > 0.01, VX_FLAG_SHADOWS, then shadow_ray with the sea-surface
 * jump, common.wgsl:1098-1148, over 16 / 48 blocks for the refract leg with / without
 * VX_FLAG_WATER_SHADOW_CUT, the frame's shadow_dist for the water reflect leg, 48 for
 * glass legs); a glass leg that ends on water is shaded by shade_water, with no sun test
 * of its own. The glass reflect leg (VX_FLAG_HI_GLASS_REFLECT) is not continued past
 * water. */
typedef struct vx_sec_hit {
    float t;            /* distance along the leg to its hit, or the leg's range */
    uint32_t voxel;     /* x | y << 6 | z << 12 in the hit chunk */
    uint32_t info;      /* chunk index | normal << 13 | VX_SEC_SUN_* | VX_SEC_TRACED | VX_SEC_HIT
                           (chunk 0x1FFF: miss) */
    uint32_t material;  /* block id at the hit */
} vx_sec_hit;
#define VX_SEC_SUN_BLOCKED (1u << 28) /* the sun ray at the hit was blocked */
#define VX_SEC_SUN_TESTED (1u << 29)  /* a sun ray ran at the hit */
#define VX_SEC_TRACED (1u << 30)
#define VX_SEC_HIT (1u << 31)
int vx_batch_set_secondary(vx_batch *b, uint32_t on);
int vx_batch_read_secondary(vx_batch *b, uint32_t view, vx_sec_hit *out, size_t n);
/* Legs queued (water legs, those behind glass included; glass legs) and sun rays run at
 * their hits in the last frame; overflow
 * is set when a queue ran out of room (never with the sizes the batch allocates). */
int vx_batch_secondary_count(vx_batch *b, uint64_t *water_legs, uint64_t *glass_legs, uint64_t *sun_rays,
                             uint32_t *overflow);

/* ---- Serving: continuous batching of view requests --------------------------------
 * No counterpart in the reference (its request loop renders and waits one view at a
 * time, src/headless.rs:419-497). Requests are copied at submit and queued in order; a
 * worker thread admits them onto free launch lanes (one HIP stream each) while view
 * slots and the HBM budget allow, and launches each lane's admitted views as one batch,
 * so requests arriving while other lanes run overlap with them on the device. Outputs
 * stay in the request's region of the device output arena until vx_serve_release.
 * Results are the bits vx_render gives for the same view with Hi-Z starting at zero.
 * The context's pools may only change (vx_pool_sync) between vx_serve_barrier, which
 * waits for every earlier request and holds later ones, and vx_serve_resume. */
typedef struct vx_serve_desc {
    uint32_t lanes;            /* concurrent batches on the device, 1..8 */
    uint32_t max_views;        /* views per batch, 1..1024 */
    uint64_t hbm_budget;       /* device bytes for in-flight request state: outputs
                                  (10 B/pixel) + inputs (chunk list, grid, tile words) */
    vx_batch_tuning tuning;    /* launch form of every batch (FLAT or PERSISTENT; variant 4 or 5,
                                  without the screen-bin bit 8: VX_E_ARG) */
} vx_serve_desc;

typedef struct vx_serve_result {
    uint64_t id;               /* the caller's request id */
    int32_t status;            /* VX_OK or the error of its batch */
    uint32_t lane, batch_views, width, height, pad;
    uint64_t batch;            /* batch sequence number (1, 2, ...) */
    double t_arrival, t_submit, t_launch, t_done;  /* seconds, vx_serve_clock */
} vx_serve_result;

typedef struct vx_serve_stats {
    uint64_t submitted, completed, launches, max_batch_views;
    uint64_t shadow_rays;      /* sun rays traced by completed batches */
    uint64_t pending, used_bytes, budget_bytes;
} vx_serve_stats;

double vx_serve_clock(void);  /* monotonic seconds, the clock of vx_serve_result */
int vx_serve_create(vx_ctx *ctx, const vx_serve_desc *d, vx_serve **out);
int vx_serve_destroy(vx_serve *s);
/* t_arrival: when the request arrived (vx_serve_clock), 0 = now. VX_E_OVERFLOW if its
 * footprint exceeds the whole budget. */
int vx_serve_submit(vx_serve *s, uint64_t id, const vx_view *view, double t_arrival);
/* Camera requests (chunk lists built on the device, as vx_batch_set_cameras builds them):
 * vx_serve_set_chunk_table uploads the LOD-0 chunk table, only while no lane runs and no
 * camera request is pending ahead of a barrier (between vx_serve_barrier and
 * vx_serve_resume, or before the first request; VX_E_BUSY otherwise); camera requests held
 * by the barrier build their lists from the new table. vx_serve_submit_camera queues a view given by its frame
 * (vxw_pack_camera_frame) and frustum planes; its results are those of vx_serve_submit
 * with the host-built list for that camera. */
int vx_serve_set_chunk_table(vx_serve *s, const vx_gpu_chunk *chunks, uint32_t n);
int vx_serve_submit_camera(vx_serve *s, uint64_t id, const vx_camera_view *view, double t_arrival);
/* Completed requests (up to cap), waiting up to wait_ms for the first. */
int vx_serve_poll(vx_serve *s, vx_serve_result *out, uint32_t cap, uint32_t *n, int wait_ms);
int vx_serve_read_vis(vx_serve *s, uint64_t id, uint64_t *out, size_t n);
int vx_serve_read_shadow(vx_serve *s, uint64_t id, uint16_t *out, size_t n);
/* Frees a completed request's outputs and budget (VX_E_BUSY while it is in flight). */
int vx_serve_release(vx_serve *s, uint64_t id);
int vx_serve_barrier(vx_serve *s);
int vx_serve_resume(vx_serve *s);
int vx_serve_stats_get(vx_serve *s, vx_serve_stats *out);
/* Native request replay (no counterpart in the reference, whose loop takes one request
 * at a time, src/headless.rs:419-497): a driver thread submits camera requests at their
 * scheduled arrival times (t0 + reqs[i].t on the vx_serve_clock, arrival-sorted; t0 0 =
 * now) and a collector thread takes their completions and releases their outputs, so no
 * caller thread sits on the arrival path. Request i gets id first_id + i. While a replay
 * runs the caller submits and polls nothing, but may apply edit batches
 * (vx_serve_barrier, vx_pool_sync, vx_serve_set_chunk_table, vx_serve_resume): requests
 * arriving during the barrier are held and render the edited world. The requests and
 * their frames are copied. One replay at a time (VX_E_BUSY). With VX_REPLAY_KEEP the
 * outputs are not released: the caller reads and releases them after the join. */
typedef struct vx_replay_req {
    double t;                  /* arrival, seconds after t0 */
    vx_camera_view view;
} vx_replay_req;
#define VX_REPLAY_KEEP 1u
int vx_serve_replay_start(vx_serve *s, const vx_replay_req *reqs, uint32_t n, uint64_t first_id, double t0,
                          uint32_t flags);
int vx_serve_replay_progress(vx_serve *s, uint32_t *submitted, uint32_t *completed);
/* Waits for every request of the replay: out[i] is request i's result (a submit error in
 * its status, with t_done 0). n must be the replay's count. */
int vx_serve_replay_join(vx_serve *s, vx_serve_result *out, uint32_t n);
const char *vx_error_string(int code);
int vx_device_name(char *buf, size_t cap);

/* ---- Device-side chunk build -------------------------------------------------------
 * build_local (src/voxel/tree.rs:89-204) for many chunks at once on the GPU: dense 64^3
 * block ids (x | z << 6 | y << 12) in, each chunk's LocalTree out as one flat record
 * (vx_local_tree header, then n_l1 vx_local_l1, n_leaves u64 leaf masks, n_attrs u32
 * attribute entries, n_palette u32 palette words; `bytes` is the whole record, a
 * multiple of 8). The records are what the host's build_local gives for the same dense
 * chunk (vxw_build_local_flat) and go into a world with vxw_insert_local_flat, which
 * interns them as insert_local does (src/voxel/world.rs:63-153). No counterpart call in
 * the reference, whose chunk builds run on CPU threads (src/stream.rs:119-146). */
typedef struct vx_local_tree {
    uint64_t root_mask, dry_mask;
    uint32_t n_l1, n_leaves, n_attrs, n_palette;
    uint32_t uniform;          /* the block id when flags & VX_LT_UNIFORM */
    uint32_t flags;            /* VX_LT_* */
    uint32_t emitters, solid_count, leaf_count;
    uint8_t aabb_min[3], aabb_max[3], pad[2];
    uint32_t bytes;
} vx_local_tree;
typedef struct vx_local_l1 {
    uint64_t mask;
    uint32_t leaf_start, leaf_prefix;
    uint32_t full, water_line;
} vx_local_l1;
#define VX_LT_UNIFORM 1u
#define VX_LT_WATER 2u
#define VX_LT_FOLIAGE 4u
#define VX_LT_CUTOUT 8u
#define VX_LT_GLASS 16u
#define VX_LT_EMITTER 32u
/* Largest record: 64 L1 nodes, 4096 leaves, 4096 attribute entries, 49 palette words per leaf. */
#define VX_LOCAL_TREE_MAX (64u + 64u * 24u + 4096u * 8u + 4096u * 4u + 4096u * 49u * 4u)
/* dense: n chunks of 64^3 u16 in host memory; out: n slots of slot_bytes (at least
 * VX_LOCAL_TREE_MAX) in host memory, record i at out + i * slot_bytes. */
int vx_build_local(vx_ctx *ctx, const uint16_t *dense, uint32_t n, void *out, size_t slot_bytes);
/* The same on device memory (dense and out on the context's device, 8-byte aligned),
 * without copies: for callers that keep chunks on the GPU, and to time the build alone. */
int vx_build_local_device(vx_ctx *ctx, const uint16_t *dense, uint32_t n, void *out, size_t slot_bytes);
/* ---- Device-resident world and edits on the GPU --------------------------------------
 * No counterpart call in the reference, whose edits run World::set_block on the CPU
 * (src/voxel/world.rs:288-529) and upload the dirty ranges (GpuMirror::sync). A device
 * world holds the pools in the per-chunk block layout (vxw_flat_layout) and the table of
 * every resident LOD-0 chunk (vxw_snapshot_chunks_all, empty chunks included, with their
 * solid counts), and the context reads its pools from it until it is destroyed.
 * vx_dworld_edit applies a batch of edit records on the GPU as set_block applies them one
 * by one: the edited chunks are decoded to dense ids, each voxel's edits are replayed in
 * order, the chunks are rebuilt (build_local on the device) and written as new blocks,
 * with the record fields set_block only grows (box, dry mask, water / foliage / cutout /
 * glass flags) grown and the water line of each L1 cell an edit changed set as set_block
 * sets it. Keys and shadow words rendered from the device world equal those of a host
 * world the same edits were applied to; the pool bytes differ (blocks are appended, runs
 * are not shared). Edits in chunks outside the table do nothing. A batch runs in passes
 * of up to 1024 chunks without waiting on the host; VX_E_OVERFLOW when a pass's blocks do
 * not fit in capacity_words (its chunks and those of later passes keep their old blocks;
 * chunks_rebuilt counts the placed ones). Replaced blocks stay until a compaction:
 * vx_dworld_compact rebuilds every non-empty chunk into a fresh buffer of the same
 * capacity, packed (records keep every field but their pointers), and vx_dworld_edit
 * compacts first when the free space is under 1.5 x the words in use after the last
 * compaction.
 * One difference from set_block: a chunk whose last solid voxel is removed and which gets
 * solid voxels again within one batch keeps the box of every voxel that became solid in
 * the batch (set_block restarts the box at the first voxel after the chunk went empty).
 * vx_dworld_chunks copies the records of the non-empty chunks in table order (the chunk
 * table view setup starts from); vx_dworld_extract decodes one chunk to dense ids.
 * A device world is destroyed before its context. */
typedef struct vx_dworld vx_dworld;
typedef struct vx_dworld_stats {
    uint64_t edits, effective_edits;  /* last batch: records, and those that changed a voxel */
    uint64_t used_bytes, capacity_bytes;
    uint32_t chunks_rebuilt, chunks;  /* last batch; table */
    float last_ms;                    /* GPU time of the last batch, decode to placement */
    uint32_t compactions;             /* compactions so far (vx_dworld_compact or at an edit) */
} vx_dworld_stats;
int vx_dworld_create(vx_ctx *ctx, const uint32_t *flat, size_t flat_words, size_t capacity_words,
                     const vx_gpu_chunk *table, const uint32_t *solid, uint32_t n, vx_dworld **out);
int vx_dworld_destroy(vx_dworld *w);
int vx_dworld_edit(vx_dworld *w, const vx_edit *edits, size_t n);
int vx_dworld_compact(vx_dworld *w);
int vx_dworld_chunks(vx_dworld *w, vx_gpu_chunk *out, uint32_t cap, uint32_t *n);
int vx_dworld_extract(vx_dworld *w, int32_t cx, int32_t cy, int32_t cz, uint16_t *dense);
int vx_dworld_stats_get(vx_dworld *w, vx_dworld_stats *out);

/* Device memory on the context's device for the *_device calls (no counterpart in the
 * reference): allocate, free, and synchronous copies to and from host memory. */
int vx_dev_alloc(vx_ctx *ctx, size_t bytes, void **out);
int vx_dev_free(vx_ctx *ctx, void *p);
int vx_dev_upload(vx_ctx *ctx, void *dst, const void *src, size_t bytes);
int vx_dev_download(vx_ctx *ctx, void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
