/*
 * vxhost.h -- C ABI of the host-side world store and per-view setup.
 *
 * This is the C++ mirror of the reference's CPU side of the path, so that a
 * caller without the Rust crate can produce the exact pool bytes, chunk lists
 * and frame records the GPU library consumes:
 *
 *   vxw_world_*           <- voxel::World               (src/voxel/world.rs:16-638)
 *   vxw_insert_dense      <- build_local + insert_local (src/voxel/tree.rs:89-204,
 *                                                        src/voxel/world.rs:63-153)
 *   vxw_get/set_block     <- World::get_block/set_block (src/voxel/world.rs:258-529)
 *   vxw_extract_dense     <- World::extract_dense       (src/voxel/world.rs:594-629)
 *   vxw_pool_*            <- RunPool/WordPool data() + take_dirty()
 *                            (src/voxel/geometry.rs:74-221, src/voxel/attributes.rs:6-88)
 *   vxw_build_view        <- Frustum::from_camera + Renderer::build_chunk_list
 *                            (src/math.rs:69-117, src/render/mod.rs:82-138,1447-1564)
 *   vxw_planner_*         <- lod::Planner (src/lod.rs:20-257)
 *   vxw_build_view_items  <- Renderer::build_chunk_list over a LOD draw list
 *   vxw_snapshot_*        <- (no counterpart) chunk records for view setup during edits
 *   vxw_pack_frame        <- the GpuFrame packing in Renderer::render
 *                            (src/render/mod.rs:1612-1710)
 *   vxw_build_local_flat / vxw_insert_local_flat
 *                         <- build_local / insert_local with a flat LocalTree record
 *                            (the device build's output format, vx_local_tree)
 *   vxw_flat_layout       <- (no counterpart) one pool block per chunk for vx_pool_set_flat
 *   vxw_shaft_*           <- shaft::ShaftField (src/shaft.rs:1-154)
 *   vxw_probe_*           <- probe::bake and its helpers (src/probe.rs:1-379)
 *   vxw_journal_parse     <- EditJournal::load (src/journal.rs:391-475)
 *   vxw_journal_write     <- EditJournal::save (src/journal.rs:477-508)
 *   vxw_generate          <- a seeded terrain generator with the reference's
 *                            structure (heightfield, sea, caves, trees); it is
 *                            fixture support and does not reproduce worldgen.rs bits.
 *
 * All calls are synchronous, CPU only, and return VXW_OK (0) or a negative code.
 * A vxw_world is not thread safe; distinct worlds may be used concurrently.
 */
#ifndef VXHOST_H
#define VXHOST_H

#include <stddef.h>
#include <stdint.h>

#include "vxmarch.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VXW_OK = 0,
    VXW_E_ARG = -3,
    VXW_E_MISSING = -7,
    VXW_E_JOURNAL = -9
};

typedef struct vxw_world vxw_world;

/* ChunkRecord fields a caller inspects (src/voxel/chunk.rs:73-111). */
typedef struct vxw_chunk_info {
    int32_t key[3];
    uint32_t root[4];            /* Inner as vec4<u32> */
    uint32_t attr_uniform;       /* 1 if AttrRef::Uniform */
    uint32_t attr_base;          /* uniform id or table base */
    uint32_t attr_len;
    int32_t aabb_min[3], aabb_max[3];
    uint32_t solid_count, leaf_count;
    uint32_t has_water, has_foliage, has_cutout, has_glass, has_emitter, emitters;
    uint64_t dry_mask;
    uint32_t lod;                /* ChunkKey::lod: voxels of 2^lod blocks */
    uint32_t pad;
} vxw_chunk_info;

typedef struct vxw_pool_stats {
    uint64_t live_elems, referenced_elems, live_runs, len;
} vxw_pool_stats;

vxw_world *vxw_world_new(void);
void vxw_world_free(vxw_world *w);

/* dense: 64^3 block ids indexed x | z<<6 | y<<12 (src/voxel/tree.rs:8-10). */
int vxw_insert_dense(vxw_world *w, int32_t cx, int32_t cy, int32_t cz, const uint16_t *dense);
int vxw_remove_chunk(vxw_world *w, int32_t cx, int32_t cy, int32_t cz);
uint16_t vxw_get_block(const vxw_world *w, int32_t x, int32_t y, int32_t z);
/* 1 if the chunk is resident (the edit applied or was a no-op), 0 otherwise. */
int vxw_set_block(vxw_world *w, int32_t x, int32_t y, int32_t z, uint16_t id);
/* Apply n 16-byte edit records (journal Record, src/journal.rs:116-122) in order.
 * Returns the number of records that landed in a resident chunk. */
int64_t vxw_apply_edits(vxw_world *w, const vx_edit *edits, size_t n);
int vxw_extract_dense(const vxw_world *w, int32_t cx, int32_t cy, int32_t cz, uint16_t *out);
int vxw_chunk_info_get(const vxw_world *w, int32_t cx, int32_t cy, int32_t cz, vxw_chunk_info *out);
/* Resident chunks of every LOD. */
size_t vxw_chunk_count(const vxw_world *w);
/* Resident LOD-0 chunk keys in insertion order (the render set order); returns their count. */
size_t vxw_chunk_keys(const vxw_world *w, int32_t *keys_xyz, size_t cap);
uint32_t vxw_emitter_chunks(const vxw_world *w);

/* ---- chunks of any LOD (ChunkKey, src/voxel/chunk.rs:6-60) ----
 * A key {pos, lod} is the chunk of 64^3 voxels of 2^lod blocks whose origin is
 * pos * (64 << lod). Block edits (vxw_set_block, vxw_apply_edits) and vxw_get_block
 * address LOD 0; a coarse chunk is inserted whole. */
#define VXW_MAX_LOD 12u
typedef struct vxw_key {
    int32_t pos[3];
    uint32_t lod;
} vxw_key;

int vxw_insert_dense_key(vxw_world *w, const vxw_key *key, const uint16_t *dense);
int vxw_remove_chunk_key(vxw_world *w, const vxw_key *key);
int vxw_extract_dense_key(const vxw_world *w, const vxw_key *key, uint16_t *out);
int vxw_chunk_info_key(const vxw_world *w, const vxw_key *key, vxw_chunk_info *out);
/* Resident chunk keys of every LOD in insertion order; returns the count. */
size_t vxw_chunk_keys_all(const vxw_world *w, vxw_key *out, size_t cap);

/* Pools: element sizes 16 (inners), 8 (leaves), 4 (bricks). */
const void *vxw_pool_data(const vxw_world *w, int pool);
size_t vxw_pool_bytes(const vxw_world *w, int pool);
/* Moves up to cap pending dirty ranges (element units) into out and clears them.
 * With out == NULL returns the pending count without clearing. */
size_t vxw_take_dirty(vxw_world *w, int pool, vx_range *out, size_t cap);
int vxw_pool_stats_get(const vxw_world *w, int pool, vxw_pool_stats *out);

/* ---- terrain generator (fixture support) ---- */
typedef struct vxw_gen_params {
    int32_t seed;
    int32_t sea_level;          /* water fills air columns below this y */
    int32_t chunk_min[3];       /* inclusive chunk key range to generate */
    int32_t chunk_max[3];       /* exclusive */
    float cave_fraction;        /* target carved fraction of the subsurface, 0 = none */
    float tree_density;         /* trees per column */
    float glass_fraction;       /* fraction of exposed surface voxels turned to glass */
    float leaf_extra;           /* fraction of exposed surface voxels turned to leaves */
    int32_t threads;            /* <= 0: hardware threads */
} vxw_gen_params;

int vxw_generate(vxw_world *w, const vxw_gen_params *p);
/* Surface height (first air y) of column (x, z) for these parameters. */
int32_t vxw_height(const vxw_gen_params *p, int32_t x, int32_t z);
/* Generates one chunk's dense grid (x | z<<6 | y<<12) without inserting it. */
int vxw_generate_dense(const vxw_gen_params *p, int32_t cx, int32_t cy, int32_t cz, uint16_t *out);
/* The chunk of any LOD: LOD 0 as vxw_generate_dense; a coarse key gets a stand-in with
 * the structure of WorldGen::generate_coarse (src/worldgen.rs:535-562): one height sample
 * per coarse column, each voxel the block at the top of its span. */
int vxw_generate_dense_key(const vxw_gen_params *p, const vxw_key *key, uint16_t *out);

/* ---- distant terrain shadows: the shaft height window (src/shaft.rs:1-154) ----
 * A VXW_SHAFT_DIM^2 window of column heights (row z, then x), one texel per texel x texel
 * blocks, centred on the camera column; vx_shaft_upload takes its heights and
 * vx_shaft_scan scans them towards the sun for the resolve's terrain_shade_beyond. The
 * frame carries the window: shaft_origin = vxw_shaft_origin_world, shaft_texel, and
 * shaft_soft = VXW_SHAFT_SOFT. */
#define VXW_SHAFT_DIM 512u    /* shaft::DIM */
#define VXW_SHAFT_TEXEL 4     /* shaft::TEXEL, the config default (src/config.rs:310) */
#define VXW_SHAFT_STEPS 8u    /* shaft::STEPS */
#define VXW_SHAFT_SOFT 4.0f   /* shaft::SOFT */
typedef struct vxw_shaft vxw_shaft;
/* ShaftField::with_texel; NULL for a texel outside 1..64. */
vxw_shaft *vxw_shaft_new(int32_t texel);
void vxw_shaft_free(vxw_shaft *s);
/* ShaftField::update with this generator's column heights (vxw_height); canopy_lift > 0
 * (tree-line canopies, off by default) is refused with VXW_E_ARG. */
int vxw_shaft_update(vxw_shaft *s, const vxw_gen_params *gen, float cam_x, float cam_z, float canopy_lift);
const float *vxw_shaft_heights(const vxw_shaft *s);
void vxw_shaft_origin_world(const vxw_shaft *s, float out[2]);
int32_t vxw_shaft_texel(const vxw_shaft *s);
/* 1 when the last update refilled the window (the heights need uploading). */
int vxw_shaft_dirty(const vxw_shaft *s);
/* ShaftField::envelope_reference: the brute-force horizon height at texel (tx, tz). */
float vxw_shaft_envelope(const vxw_shaft *s, const float sun[3], uint32_t tx, uint32_t tz);

/* ---- ambient probes: the bake (src/probe.rs:1-379) ----
 * A chunk's 16^3 probes (4 blocks apart, from origin + 2): per cube axis (-X, +X, -Y, +Y,
 * -Z, +Z) the sky occlusion occl[axis * 4096 + (pz * 16 + py) * 16 + px] and the ground
 * bounce colour bounce[(axis * 4096 + i) * 3 + c], from this generator's column heights
 * and surface blocks (vxw_height, vxw_surface_block) along 16 azimuths to 192 blocks;
 * bounce_shadow weights each ground sample by its own sky exposure. vx_probe_write
 * uploads the result. */
#define VXW_PROBES 4096u
int vxw_probe_bake(const vxw_gen_params *gen, const int32_t origin[3], int bounce_shadow, float *occl,
                   float *bounce);
/* probe::cube_from_tangents: the occlusion cube of one probe from its 16 horizon tangents. */
void vxw_probe_cube_from_tangents(const float tan_h[16], float out[6]);
/* probe::bounce_of_uniform_ground: the bounce of ground of one albedo all round (6 x rgb). */
void vxw_probe_bounce_of_uniform_ground(const float albedo[3], float out[18]);
/* probe::uploadable: the chunk lies within UPLOAD_REACH (224 blocks) of the camera in xz. */
int vxw_probe_uploadable(const int32_t origin[3], const float cam[3]);
/* The block at the top of column (x, z) of height h (the terrain, or the sea surface over it). */
uint16_t vxw_surface_block(const vxw_gen_params *p, int32_t x, int32_t z, int32_t h);
/* block::ALBEDO[id][c] (src/block.rs:187-214). */
float vxw_albedo(uint16_t id, int c);

/* ---- per-view setup ---- */
typedef struct vxw_camera {
    float pos[3];       /* eye position */
    float fwd[3], right[3], up[3];
    float fov_y;        /* radians */
    uint32_t width, height;
    float far_;         /* view distance */
    float sun_dir[3];
    float daylight;
    float sea_level;
    float time;
    float jitter[2];
    uint32_t flags;     /* FLAG_* word (src/render/mod.rs:144-270) */
    uint32_t frame_index;
    float waves[4];     /* wave amp, scale, speed, reflect slope (config.rs defaults 0.12, 32, 0.32, 0) */
} vxw_camera;

/* Player::look_dir + scene::camera_basis (src/player.rs:122-126, src/scene.rs:334-339). */
void vxw_camera_basis(float yaw, float pitch, float fwd[3], float right[3], float up[3]);
/* scene::sun (src/scene.rs:10-15). */
void vxw_sun(float time_of_day, float dir[3], float *daylight);

typedef struct vxw_view {
    uint32_t n_visible;     /* frame.chunk_count */
    uint32_t n_total;       /* visible + off-screen shadow casters */
    uint32_t camera_chunk;  /* index among visible, or 0xFFFFFFFF */
    int32_t grid_min[3];
} vxw_view;

/* Builds the distance-sorted chunk list and the 32x8x32 chunk grid for a camera.
 * chunks must hold VX_MAX_CHUNKS entries, grid VX_GRID_LEN. */
int vxw_build_view(const vxw_world *w, const vxw_camera *cam, vx_gpu_chunk *chunks, uint32_t *grid,
                   vxw_view *out);
/* View snapshots (no counterpart in the reference, whose one owner thread edits and
 * renders in turn): a copy of the LOD-0 chunk records vxw_build_view reads, so that a
 * server can set up views from the world as it was while an edit batch is applied to
 * it. vxw_build_view_snapshot gives what vxw_build_view gave at vxw_snapshot_take.
 * A snapshot is read-only between takes and may be read from several threads. */
typedef struct vxw_snapshot vxw_snapshot;
vxw_snapshot *vxw_snapshot_new(void);
void vxw_snapshot_free(vxw_snapshot *s);
int vxw_snapshot_take(vxw_snapshot *s, const vxw_world *w);
uint32_t vxw_snapshot_emitter_chunks(const vxw_snapshot *s);
int vxw_build_view_snapshot(const vxw_snapshot *s, const vxw_camera *cam, vx_gpu_chunk *chunks, uint32_t *grid,
                            vxw_view *out);
/* The chunk table the device builds views from (vx_batch_set_chunk_table): the GpuChunk
 * record (fade 1) of every snapshot chunk with solid voxels, in the order vxw_build_view
 * takes them; returns the count (out may be NULL). */
size_t vxw_snapshot_chunks(const vxw_snapshot *s, vx_gpu_chunk *out, size_t cap);
/* Every resident LOD-0 chunk in world order, the empty ones included, with its solid
 * voxel count (ChunkRecord::solid_count): the table a device world starts from
 * (vx_dworld_create). */
size_t vxw_snapshot_chunks_all(const vxw_snapshot *s, vx_gpu_chunk *out, uint32_t *solid, size_t cap);
/* The three pools re-laid out as one contiguous, 128-byte aligned block per chunk
 * (no counterpart in the reference, whose pools share interned runs between chunks,
 * src/voxel/geometry.rs:12-52, attributes.rs:6-88): the chunk's L1 nodes, the leaf runs
 * of its L1 cells, its attribute table and the palettes that table points to, with
 * every pointer rewritten into the block. chunks_out receives the n records with their
 * root run pointer and attr_base moved; all other fields are copied. Element counts:
 * n_inners in 16-byte nodes, n_leaves in 8-byte leaves, n_bricks in words. flat == NULL
 * only sizes the buffer (*words). Upload the words with vx_pool_set_flat and render
 * with the returned records. Chunks with light tables (light_flags bit 0 clear) are
 * refused: only uniform light (LightRef::Uniform) is carried over. */
int vxw_flat_layout(const uint32_t *inners, size_t n_inners, const uint32_t *leaves, size_t n_leaves,
                    const uint32_t *bricks, size_t n_bricks, const vx_gpu_chunk *chunks, uint32_t n,
                    vx_gpu_chunk *chunks_out, uint32_t *flat, size_t cap_words, size_t *words);
/* Frustum::from_camera (src/math.rs:69-117) of a camera: six planes (nx, ny, nz, d); a
 * point p is outside plane i when nx*px + ny*py + nz*pz + d < 0. */
int vxw_frustum_planes(const vxw_camera *cam, float planes[24]);
/* The GpuFrame of a camera whose chunk list the device builds (vx_batch_set_cameras):
 * vxw_pack_frame with the camera's grid origin, no chunks counted yet. */
int vxw_pack_camera_frame(const vxw_camera *cam, uint32_t max_pairs, vx_gpu_frame *out);
/* Packs the GpuFrame record for a camera and view (max_pairs 0 = pairs_for(tiles)). */
int vxw_pack_frame(const vxw_camera *cam, const vxw_view *view, uint32_t max_pairs, vx_gpu_frame *out);
/* render::pairs_for (src/render/mod.rs:27-30). */
uint32_t vxw_pairs_for(uint32_t tiles_x, uint32_t tiles_y);

/* ---- LOD planning (src/lod.rs:20-257) ----
 * vxw_planner_plan <- Planner::plan (src/lod.rs:108-152): the draw list of a camera
 * (chunks of several LODs; a chunk changing LOD draws beside its children with fade -f
 * against their f) and the keys it wants resident. Residency comes from the world
 * (w != NULL: absent, empty = no solid voxel, solid), else from fn(user, key), else
 * every key is solid. A planner keeps the set of split keys between calls (the
 * hysteresis of the no-cross-fade mode). */
enum { VXW_FADE_SMOOTHSTEP = 0, VXW_FADE_SMOOTHERSTEP = 1 };
enum { VXW_ABSENT = 0, VXW_EMPTY = 1, VXW_SOLID = 2 };

typedef struct vxw_lod_config {      /* LodConfig (src/lod.rs:20-53) */
    uint32_t max_lod;                /* 4 */
    float factor;                    /* 2.0: a key splits inside size * factor of its centre */
    float view_distance;             /* 2600 */
    float fade_band;                 /* 1.25: the fade runs to size * factor * fade_band */
    uint32_t fade_schedule;          /* VXW_FADE_* */
    uint32_t cover_depth;            /* 2 */
    uint32_t cross_fade;             /* 1 */
    uint32_t shadow_share;           /* 1: the chunk grid prefers the larger share */
    uint32_t offscreen_shadows;      /* 1: off-screen grid chunks cast shadows */
} vxw_lod_config;

typedef struct vxw_render_item {     /* RenderItem (src/lod.rs:55-81) */
    vxw_key key;
    float fade;                      /* 1 solid; f in (0,1) a fading child; -f its parent */
} vxw_render_item;

typedef int (*vxw_residency_fn)(void *user, const vxw_key *key);
typedef struct vxw_planner vxw_planner;

void vxw_lod_config_default(vxw_lod_config *cfg);
vxw_planner *vxw_planner_new(void);
void vxw_planner_free(vxw_planner *p);
int vxw_planner_plan(vxw_planner *p, const float cam[3], const vxw_lod_config *cfg, const vxw_world *w,
                     vxw_residency_fn fn, void *user);
/* The last plan's draw list / wanted keys; return the count (out may be NULL). */
size_t vxw_planner_render(const vxw_planner *p, vxw_render_item *out, size_t cap);
size_t vxw_planner_desired(const vxw_planner *p, vxw_key *out, size_t cap);

/* Renderer::build_chunk_list (src/render/mod.rs:1447-1564) over a draw list: items not
 * resident or without solid voxels are dropped; GpuChunk voxel_size 2^lod, origin and
 * fade from the item; the grid is filled in (share, -lod) order when cfg->shadow_share
 * (cfg NULL: vxw_lod_config_default). vxw_build_view is this over every LOD-0 chunk in
 * insertion order with fade 1. */
int vxw_build_view_items(const vxw_world *w, const vxw_camera *cam, const vxw_render_item *items, size_t n,
                         const vxw_lod_config *cfg, vx_gpu_chunk *chunks, uint32_t *grid, vxw_view *out);

/* build_local (src/voxel/tree.rs:89-204) of one dense chunk as a flat vx_local_tree record
 * (the format vx_build_local writes); *written = record bytes (out may be NULL to learn
 * the size, at most VX_LOCAL_TREE_MAX). */
int vxw_build_local_flat(const uint16_t *dense, void *out, size_t cap, size_t *written);
/* insert_local (src/voxel/world.rs:63-153) of a flat vx_local_tree record. */
int vxw_insert_local_flat(vxw_world *w, int32_t cx, int32_t cy, int32_t cz, const void *rec, size_t bytes);

/* ---- Edit journal files (VXJ1, src/journal.rs:13-139) -------------------------------
 * Header (24 B) | PlayerState (32 B, flags & 1) | BarState (24 B, flags & 4) | count
 * journal records (vx_edit, 16 B). flags & 2 marks an append log: records past the
 * header's count are recovered, a torn trailing record dropped. */
typedef struct vxw_player {
    float pos[3];
    float yaw, pitch;
    uint32_t hotbar, fly, pad;
} vxw_player;

typedef struct vxw_journal_info {
    uint32_t version, flags;
    int32_t seed, sea_level;
    uint32_t count;              /* edits in the file (after append recovery) */
    uint32_t recovered;          /* of which past the header's count (append log) */
    uint32_t torn_bytes;         /* trailing bytes of a half-written record, dropped */
    uint32_t has_player, has_bar;
    vxw_player player;
    uint16_t bar[10];
} vxw_journal_info;

/* Checks the file as EditJournal::load does (magic, version 1..=3, seed and sea level,
 * player and bar blocks, record count) and copies its records into edits (cap >= count;
 * edits may be NULL to learn the count). VXW_E_JOURNAL with the reason in err. Records
 * come in file order: applied in order (vxw_apply_edits) the last edit of a voxel wins,
 * as in the journal's edit map. */
int vxw_journal_parse(const void *bytes, size_t n, int32_t seed, int32_t sea_level, vxw_journal_info *info,
                      vx_edit *edits, size_t cap, char *err, size_t err_cap);
/* Writes the journal EditJournal::save writes: the last edit of each voxel, sorted by
 * (x, z, y), with the player / bar blocks when given (NULL: absent). *written = bytes
 * needed; out may be NULL to learn it. */
int vxw_journal_write(const vx_edit *edits, size_t n, int32_t seed, int32_t sea_level, const vxw_player *player,
                      const uint16_t *bar, void *out, size_t cap, size_t *written);

#ifdef __cplusplus
}
#endif
#endif
