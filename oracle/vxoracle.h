/*
 * vxoracle.h -- CPU oracle for the voxel ray-march hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker: a plain-C restatement of the
 * reference's WGSL passes (primary march, Hi-Z tile culling, sun shadow) in the
 * reference's f32 operation order, compiled with -ffp-contract=off. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product library never links or calls it.
 *
 * What pins it (tests/): the reference's literal tests (key packing, pair capacity,
 * tree round trips, ray_dir against view_proj, the journal file); float64 brute force
 * (the single-chunk march at voxel sizes 1-8, whole frames over multi-chunk worlds: first
 * hit, face, distance and the primary sun bit; leaf cutouts at fill 1 and 0; the
 * secondary legs and their sun tests, including the sea-surface jump); closed forms
 * (the heatmap counts on full chunks, the fade dither's exclusivity); invariants (Hi-Z
 * on = off, thread decomposition).
 * Also pinned on whole frames: water surfaces, glass and glowstone hits from above the
 * water, rays under water skipping water (with the water-line skip), and sun rays
 * passing water, glass and (emitter worlds) glowstone.
 * Leaf cutouts at the fills the flags give (0.62, 0.72): a float64 DDA at micro
 * resolution with an integer form of the hash threshold, on whole frames (keys, micro
 * cells, sun bits), on random rays into a chunk at negative coordinates, and on micro
 * cells whose hash lies exactly on the threshold (found through the hash's inverse). The
 * water-interface probe: rays crossing from a neighbouring chunk's water, against a
 * float64 DDA whose stop rule is "water only when entered from a non-water cell".
 * Planted mistakes (VXO_PLANT in vxoracle.c) in either are caught by those checks.
 * The wave field and the waved water reflect direction: cos and exp bounded against
 * float64; the field as the gradient of the float64 sum of its octaves' heights, and
 * curl-free; the unresolved variance in closed form (isotropic ends, anisotropic fade);
 * the reflect direction against a float64 construction over deep and shallow water
 * (shoal); planted mistakes in them are caught. The water legs behind glass: the same
 * float64 leg construction from the point the transmit leg meets the water.
 * The Snell bend of underwater rays: a float64 refraction at the sea surface and DDA
 * through the chunk, with planted mistakes caught. Foliage tufts and the water-look
 * ripples are not restated; the library refuses them.
 *
 * Every struct here is an independent restatement of the reference layouts:
 *   vxo_frame  <- GpuFrame  (src/render/mod.rs:398-466, 352 bytes)
 *   vxo_chunk  <- GpuChunk  (src/render/mod.rs:468-489, 96 bytes)
 * Pools are the reference's three storage buffers
 *   inners: array<vec4<u32>>, leaves: array<vec2<u32>>, bricks: array<u32>
 *   (src/render/shaders/common.wgsl:100-104).
 */
#ifndef VXORACLE_H
#define VXORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vxo_frame {
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
    float wave_amp, wave_scale, wave_speed, wave_reflect_slope;  /* offsets 240-255 */
    float shaft_origin[2], shaft_texel, shaft_soft, cloud_patch, cloud_relief;  /* offset 256 */
    float haze_warm, zenith_deep;
    float prev_view_proj[16];  /* read by vxo_taa only */
} vxo_frame;

typedef struct vxo_chunk {
    float aabb_min[3];  float voxel_size;
    float aabb_max[3];  uint32_t lod;
    float origin[3];    float fade;
    uint32_t root[4];
    uint32_t attr_base, attr_flags, light_base, light_flags;
    uint32_t dry_mask[2];
    uint32_t reserved[2];
} vxo_chunk;

typedef struct vxo_pools {
    const uint32_t *inners;  size_t n_inners;   /* count of vec4<u32> */
    const uint32_t *leaves;  size_t n_leaves;   /* count of vec2<u32> */
    const uint32_t *bricks;  size_t n_bricks;   /* count of u32 */
} vxo_pools;

/* WGSL override constants the march reads (src/render/shaders/common.wgsl:180-263,
 * set per pipeline in src/render/mod.rs:2355-2689). */
typedef struct vxo_spec {
    uint32_t leaf_cutout;     /* SPEC_LEAF_CUTOUT */
    uint32_t glass;           /* SPEC_GLASS */
    uint32_t full_march;      /* SPEC_FULL_MARCH */
    uint32_t emitter_gather;  /* SPEC_EMITTER_GATHER */
    uint32_t water_repair;    /* SPEC_WATER_REPAIR */
    float leaf_fill;          /* SPEC_LEAF_FILL */
    uint32_t glass_reflect;   /* SPEC_GLASS_REFLECT */
    uint32_t waves, wave_aniso, wave_shoal, wave_fill;  /* SPEC_WAVES, SPEC_WAVE_ANISO/SHOAL/FILL */
    uint32_t snell_bend;       /* SPEC_SNELL_BEND (march.wgsl:23-45) */
} vxo_spec;

/* Algorithmic traffic and work of one frame, counted exactly by the restatement. */
typedef struct vxo_stats {
    uint64_t pairs_select;     /* counters[4]: raw tile_select emits */
    uint64_t pairs_deferred;   /* counters[1]: raw Hi-Z deferrals */
    uint64_t pairs_recover;    /* counters[5]: raw recovered pairs */
    uint64_t rays_primary;     /* (pixel, pair) marches issued */
    uint64_t rays_shadow;      /* pixels whose shadow_guide is valid */
    uint64_t chunk_marches;    /* march_chunk calls (primary + shadow) */
    uint64_t iters;            /* DDA iterations */
    uint64_t l1_loads, leaf_loads, attr_loads, brick_words, grid_loads, chunk_loads;
    uint64_t node_bytes;       /* 96*chunk + 16*l1 + 8*leaf + 4*(attr+brick+grid) */
    uint64_t capped;           /* marches stopped by MAX_ITERS */
    uint64_t oob;              /* pool reads outside the supplied arrays */
    uint64_t rays_secondary;   /* secondary legs traced (trace_world calls) */
    uint64_t rays_secondary_sun; /* shadow_ray calls at secondary hits */
} vxo_stats;

/* Single-function restatements exposed for pinning tests. */
void     vxo_ray_dir(const vxo_frame *f, float px, float py, float out[3]);
uint32_t vxo_depth_key(const vxo_frame *f, float t);
uint64_t vxo_make_key(const vxo_frame *f, float t, uint32_t normal, uint32_t ci,
                      const uint32_t v[3], const uint32_t micro[3]);
uint32_t vxo_hash_u32(uint32_t x);
uint16_t vxo_f32_to_f16(float x);
uint32_t vxo_block_at(const vxo_pools *p, const vxo_chunk *chunks, uint32_t ci,
                      const uint32_t v[3]);

typedef struct vxo_hit {
    uint32_t hit;
    float t;
    uint32_t normal;
    uint32_t voxel[3];
    uint32_t micro[3];
    uint32_t iters;
} vxo_hit;

void vxo_march_chunk(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks,
                     const uint32_t *grid, const vxo_frame *f, const vxo_spec *s,
                     uint32_t ci, const float ro_w[3], const float rd[3], float t_limit,
                     int skip_water, int skip_foliage, int skip_glass, int sun_ray,
                     vxo_hit *out);

/* cos and exp as the wave field evaluates them (the same f32 steps as the HIP kernels;
 * WGSL leaves their accuracy to the implementation). */
float vxo_wave_cos(float x);
float vxo_wave_exp(float x);
float vxo_wave_sin(float x);
/* wave_field (resolve.wgsl:946-967): x, y = slope gradient, z = unresolved variance. */
void vxo_wave_field(const vxo_frame *f, const vxo_spec *s, float px, float pz, float px_world,
                    const float vh[2], float kk, int aniso, float shoal, float out[3]);

/* water_ctx's reflect-leg direction mirror_t at a water hit (resolve.wgsl:1078-1167). */
void vxo_water_reflect_dir(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const uint32_t *grid,
                           const vxo_frame *f, const vxo_spec *s, uint32_t nid, const float hit[3],
                           const float rd[3], float t, float out[3]);

float vxo_shadow_ray_t(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks,
                       const uint32_t *grid, const vxo_frame *f, const vxo_spec *s,
                       const float ro[3], const float rd[3], float max_dist);

/*
 * One frame of the reference pass sequence (src/render/mod.rs:1716-1795,1821-1840):
 * clear -> tile_select -> march -> [build_hiz -> recover_select -> march] -> primary_shadow.
 * hiz: in = previous frame's Hi-Z (T entries; ignored and zeroed when FLAG_HIZ is off),
 *      out = this frame's Hi-Z as the reference leaves it (built after the first march).
 * The pair list is never truncated here; the raw demand is reported in stats so a
 * caller can check it against frame->max_pairs.
 * vis: W*H keys. shadow: W*H f16 bit patterns (R16F target). Either may be NULL to skip.
 * threads <= 0 selects the hardware thread count.
 * Returns 0, or -1 on bad arguments.
 */
int vxo_render(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks,
               const uint32_t *grid, const vxo_frame *f, const vxo_spec *s,
               uint32_t *hiz, uint64_t *vis, uint16_t *shadow, vxo_stats *stats,
               int threads);

/* vxo_render with the FLAG_HEATMAP counter: heat[W*H] receives, for a frame whose flags
 * hold FLAG_HEATMAP, the sum of march_chunk iterations over every pair marched at each
 * pixel (march.wgsl:51-53), zero elsewhere; heat may be NULL. */
int vxo_render_heat(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks,
                    const uint32_t *grid, const vxo_frame *f, const vxo_spec *s,
                    uint32_t *hiz, uint64_t *vis, uint16_t *shadow, uint32_t *heat,
                    vxo_stats *stats, int threads);

/* One secondary leg: trace_world's result (common.wgsl:1205-1253) plus the block id at its
 * hit. info = ci | normal << 13 | traced << 30 | hit << 31 (ci = 0x1FFF on a miss). */
typedef struct vxo_sec {
    float t;            /* hit distance along the leg, or its max_dist */
    uint32_t voxel;     /* x | y << 6 | z << 12 */
    uint32_t info;      /* ci | normal << 13 | sun blocked << 28 | sun tested << 29 | traced << 30 | hit << 31 */
    uint32_t material;  /* block_at at the hit, 0 otherwise */
} vxo_sec;

/* Secondary bounce of every pixel of a visibility buffer (the legs of resolve.wgsl's
 * shade_water and shade_glass): out[4*pix] is the refract (water) or transmit (glass) leg,
 * out[4*pix+1] the reflect leg (about water_ctx's wave normal on waved water), and, when a
 * glass transmit leg ends on water, out[4*pix+2] and out[4*pix+3] are the refract and
 * reflect legs shade_water traces from that point. */
int vxo_secondary(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const uint32_t *grid,
                  const vxo_frame *f, const vxo_spec *s, const uint64_t *vis, vxo_sec *out,
                  vxo_stats *stats, int threads);

/* Shadow pass alone, from a given visibility buffer. */
int vxo_primary_shadow(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks,
                       const uint32_t *grid, const vxo_frame *f, const vxo_spec *s,
                       const uint64_t *vis, uint16_t *shadow, vxo_stats *stats, int threads);

/* Decode one visibility key into world voxel, face, material and f32 distance
 * (shadow_guide/hit_t, src/render/shaders/shadow.wgsl:8-30, common.wgsl:1300-1318). */
typedef struct vxo_decoded {
    int32_t voxel[3];     /* world voxel coordinate (origin + v*vs) */
    uint32_t face;        /* normal id 0..7 */
    uint32_t material;    /* block id */
    uint32_t ci;
    float t;              /* hit distance re-derived by hit_t */
} vxo_decoded;

int vxo_decode(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks,
               const vxo_frame *f, const uint64_t *vis, vxo_decoded *out, int threads);

int vxo_hardware_threads(void);

/* ---- Resolve, lane 0 (resolve.wgsl:1378-1466), restated in vxresolve.inc ---- */
/* log2 / exp2 as the library evaluates them; pow(x, y) = exp2(y * log2(x)) (WGSL). */
float vxo_log2(float x);
float vxo_exp2(float x);
float vxo_pow(float x, float y);

/* The override switches the resolve reads, from the pipeline key (src/render/mod.rs:
 * 2355-2689): SPEC_SKY_SPECULAR, SPEC_LIGHTING_REPAIR, SPEC_SKY_TINT, SPEC_SKY_COOL,
 * SPEC_SKY_LOOK, SPEC_LIGHT_RGB, SPEC_EMITTER_GATHER, SPEC_SHORE_WET,
 * SPEC_TEX_VARIATION, SPEC_GLASS. */
typedef struct vxo_rspec {
    uint32_t sky_specular, lighting_repair, sky_tint, sky_cool, sky_look, light_rgb, emitter_gather,
        shore_wet, tex_variation, glass;
    /* Lanes 1 and 2 (glass and water pixels; 0: lane 0 only, those pixels stay at alpha
     * 0), SPEC_WATER_SEC with WATER_SEC_SHIFT 1, SPEC_SHORE_FOAM, SPEC_FLAT_SECONDARY, and the
     * march switches the legs' traces and water_ctx read (waves, water repair, glass
     * reflect); march may be NULL when lanes is 0. */
    uint32_t lanes, water_sec, shore_foam, flat_secondary;
    const vxo_spec *march;
    /* SPEC_DISTANT_SHADOWS: terrain_shade_beyond from the shaft field's result slice
     * (VXO_SHAFT_DIM^2 f32, the output of vxo_shaft_scan), read when the frame has
     * FLAG_DISTANT_SHADOWS; shaft may be NULL when distant is 0. */
    uint32_t distant;
    const float *shaft;
    /* SPEC_TINT (the biome tint where the frame has FLAG_TINT), SPEC_TINT_BALANCE */
    uint32_t tint, tint_balance;
    /* SPEC_PROBE_CUBE, _BOUNCE, _SUN, _SUN_HIGH over the probe field `probe` (128 x 768 x
     * 128 RGBA16F texels, texel (x, y, z) at ((z * 768 + y) * 128 + x) * 4; NULL when
     * neither cube nor bounce is on) */
    uint32_t probe_cube, probe_bounce, probe_sun, probe_sun_high;
    const uint16_t *probe;
    uint32_t caustics;  /* SPEC_CAUSTICS */
    /* SPEC_SOFT_SHADOWS: lane-0 pixels take the sun visibility from the shadow_horizontal /
     * shadow_vertical mask (R8) built over vis and sun first; sun must not be NULL */
    uint32_t soft_shadows;
    /* SPEC_PROBE_TAP (the gathered ambient times probe_tap) and SPEC_PROBE_AMBIENT (probe_tap
     * alone, face shade 1, no floor light); both read `probe` */
    uint32_t probe_tap, probe_ambient;
} vxo_rspec;

/* Distant terrain shadows (src/shaft.rs, shaft.wgsl:1-24, common.wgsl:1557-1580).
 * vxo_shaft_scan: the SHAFT_STEPS log-step passes of shaft_scan over a VXO_SHAFT_DIM^2
 * height field with the frame's sun_dir and shaft_texel; result = the slice pass 7
 * writes (SHAFT_RESULT). vxo_terrain_shade: terrain_shade(p + sun_dir * d) over such a
 * result with the frame's shaft_origin, shaft_texel and shaft_soft (d = 0: terrain_shade). */
#define VXO_SHAFT_DIM 512u
#define VXO_SHAFT_STEPS 8u
int vxo_shaft_scan(const float *heights, const vxo_frame *f, float *result);
/* The biome tint (resolve.wgsl:166-287): biome_noise (2D OpenSimplex2 as FastNoiseLite
 * 1.1.1's SingleOpenSimplex2, normalised to [-1, 1]) and biome_tint at world (x, z) with
 * the frame's tint_seed and tint frequencies (balance: TINT_PLAINS_BALANCED). */
float vxo_biome_noise(int32_t seed, float freq, float x, float z);
/* probe_field (resolve.wgsl:106-118) at world p for face normal nid: rgba. */
void vxo_probe_field(const uint16_t *probe, const float p[3], uint32_t nid, float out[4]);
/* probe_tap (resolve.wgsl:120-127) at world p: rgb. */
void vxo_probe_tap(const uint16_t *probe, const float p[3], float out[3]);
void vxo_biome_tint(const vxo_frame *f, uint32_t balance, float x, float z, float out[3]);
/* Soft shadows alone (shadow.wgsl:49-116): shadow_horizontal's RG16F output into h_out
 * (W * H * 2 f16 words; may be NULL) and shadow_vertical's R8 mask into mask_out (W * H),
 * for the frame, the rspec's glass switch, vis and the primary_shadow words sun. */
int vxo_soft_mask(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const uint32_t *grid,
                  const vxo_frame *f, const vxo_rspec *r, const uint64_t *vis, const uint16_t *sun, uint16_t *h_out,
                  uint8_t *mask_out, int threads);
float vxo_terrain_shade(const float *result, const vxo_frame *f, const float p[3], float d);

/* The caller's atlas (RGBA8 sRGB: mip level 0 of every layer, then level 1, ...; level l
 * is (size >> l)^2 texels a layer) and block_faces table (8 face words per block id:
 * layer | perm class << 16, src/block.rs:96-102). */
typedef struct vxo_resolve_assets {
    const uint8_t *atlas;
    uint32_t size, layers, mips;
    const uint32_t *block_faces;
    uint32_t n_faces;
} vxo_resolve_assets;

/* out[4 * pix]: the lane-0 colour (rgb, alpha 1), or all zero where lane 0 writes
 * nothing (water and glass pixels). sun: the shadow pass's R16F words; heat: the
 * iteration counter of a FLAG_HEATMAP frame (either may be NULL when not read). */
int vxo_resolve(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const uint32_t *grid,
                const vxo_frame *f, const vxo_rspec *r, const vxo_resolve_assets *a, const uint64_t *vis,
                const uint16_t *sun, const uint32_t *heat, float *out, vxo_stats *stats, int threads);

/* Present (blit.wgsl:18-93): w*h RGBA32F colours to RGBA8 (alpha 255). tone 1: ACES fit,
 * else the knee tonemap; grade 1 / 2 mixed by strength below 1; encode: sRGB encode. */
/* TAA (taa.wgsl:1-127) after the resolve: src is the resolve colour (W*H*4 f32, rounded
 * to f16 as the trace output stores it), hist the previous TAA output (W*H*4 f16 bits);
 * out receives this frame's output (W*H*4 f16 bits, alpha 1). */
int vxo_taa(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const vxo_frame *f, const uint64_t *vis,
            const float *src, const uint16_t *hist, uint16_t *out);
void vxo_present(const float *rgba, uint32_t w, uint32_t h, uint32_t encode, uint32_t tone, uint32_t grade,
                 float strength, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
