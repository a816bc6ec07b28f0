/*
 * vxoracle.c -- CPU oracle (test infrastructure only; see vxoracle.h).
 *
 * A line-by-line C restatement of the reference WGSL for the three hot passes.
 * Every function cites the WGSL it follows. Arithmetic is IEEE f32 in the WGSL
 * operation order: the build compiles this file with -O2 -ffp-contract=off and
 * no fast-math, so each + - * / sqrt rounds once, exactly as written.
 *
 * Conventions taken for WGSL builtins whose rounding WGSL leaves open:
 *   length(v)    = sqrt(dot(v, v)), dot summed x, then y, then z
 *   normalize(v) = v / length(v)   (three divisions)
 *   min/max      = IEEE minNum/maxNum (fminf/fmaxf)
 *   i32(f32)     = truncation toward zero (inputs stay in range on these paths)
 *   R16F store   = round-to-nearest-even f32 -> f16
 */
#include "vxoracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ---- constants (src/render/shaders/common.wgsl:140-175, :264-296, :489-512) ---- */
#define FULL_BIT 0x80000000u
#define PTR_MASK 0x7FFFFFFFu
#define UNIFORM_BIT 0x80000000u
#define NO_CHUNK 0xFFFFFFFFu
#define FLAG_HIZ 1u
#define FLAG_HEATMAP 2u
#define FLAG_SHADOWS 4u
#define FLAG_UNDERWATER 64u
#define FLAG_WATER_FAR 2097152u
#define FLAG_WATER_DARK 4194304u
#define MAX_ITERS 512u
#define WATER_FAR_DIST 128.0f
#define WATER_DARK_DEPTH 15.0f
#define WATER_DARK_DIST 64.0f
#define WATER_ID 13u
#define GLASS_ID 10u
#define GLOWSTONE_ID 11u
#define LEAVES_ID 7u
#define PINE_LEAVES_ID 16u
#define ATTR_UNIFORM 1u
#define ATTR_HAS_WATER 2u
#define ATTR_HAS_FOLIAGE 4u
#define ATTR_HAS_CUTOUT 8u
#define ATTR_HAS_GLASS 16u
#define ATTR_HAS_EMITTER 32u
#define NORMAL_CROSS_A 6u
#define NORMAL_CROSS_B 7u
#define MICRO 4u
#define MICRO_SHIFT 6u
#define VOXEL_MASK 63u

/* Planted mistakes, for the tests that show their float64 references can tell a wrong
 * restatement from a right one (tests/test_oracle_cutout_water_pins.py compiles copies
 * of this file with -DVXO_PLANT=n into a temporary directory). 0, the only value any
 * shipped build uses, plants nothing.
 *   1: leaf_solid compares with <= instead of <
 *   2: leaf_solid hashes the voxel without its micro offset scaled (q = wv + m)
 *   3: micro_entry snaps to the far micro layer instead of the entry one
 *   4: cutout_march steps to the near bound of the micro cell instead of the far one
 *   5: the water-interface probe looks inward (probe = boundary - outward * off)
 *   6: the water-interface probe is dropped (a water voxel always stops the ray) */
#ifndef VXO_PLANT
#define VXO_PLANT 0
#endif
/*   7: wave_octave adds the phase with the wrong sign
 *   8: wave_octave swaps the components of its direction in the dot product
 *   9: the anisotropic fade takes exp(-u) instead of exp(-u * u)
 *  10: the reflect direction skips the optical-roughness blend toward the face normal */

/* ---- per-thread context: pools + counters ---- */
typedef struct ctx {
    const vxo_pools *p;
    const vxo_chunk *chunks;
    uint32_t n_chunks;
    const uint32_t *grid;
    const vxo_frame *f;
    const vxo_spec *s;
    vxo_stats st;
} ctx;

static void inner_at(ctx *c, uint32_t idx, uint32_t out[4]) {
    if ((size_t)idx >= c->p->n_inners) {
        c->st.oob++;
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    memcpy(out, c->p->inners + 4 * (size_t)idx, 16);
}

static void leaf_at(ctx *c, uint32_t idx, uint32_t out[2]) {
    if ((size_t)idx >= c->p->n_leaves) {
        c->st.oob++;
        out[0] = out[1] = 0;
        return;
    }
    memcpy(out, c->p->leaves + 2 * (size_t)idx, 8);
}

static uint32_t brick_at(ctx *c, uint32_t idx) {
    if ((size_t)idx >= c->p->n_bricks) {
        c->st.oob++;
        return 0;
    }
    return c->p->bricks[idx];
}

static const vxo_chunk *chunk_at(ctx *c, uint32_t ci) {
    static const vxo_chunk zero;
    if (ci >= c->n_chunks) {
        c->st.oob++;
        return &zero;
    }
    return &c->chunks[ci];
}

static uint32_t grid_at(ctx *c, uint32_t idx) {
    uint32_t n = c->f->grid_dim[0] * c->f->grid_dim[1] * c->f->grid_dim[2];
    c->st.grid_loads++;
    if (idx >= n) {
        c->st.oob++;
        return NO_CHUNK;
    }
    return c->grid[idx];
}

/* ---- mask helpers (common.wgsl:299-318) ---- */
static int mask_has(const uint32_t m[2], uint32_t bit) {
    if (bit < 32u) return ((m[0] >> bit) & 1u) != 0u;
    return ((m[1] >> (bit - 32u)) & 1u) != 0u;
}

static uint32_t mask_below(const uint32_t m[2], uint32_t bit) {
    if (bit < 32u) return (uint32_t)__builtin_popcount(m[0] & ((1u << bit) - 1u));
    if (bit == 32u) return (uint32_t)__builtin_popcount(m[0]);
    return (uint32_t)__builtin_popcount(m[0]) +
           (uint32_t)__builtin_popcount(m[1] & ((1u << (bit - 32u)) - 1u));
}

static uint32_t cell_bit(uint32_t x, uint32_t y, uint32_t z) { return x | (y << 2u) | (z << 4u); }

/* ---- attributes (common.wgsl:320-356) ---- */
static uint32_t leaf_block(ctx *c, uint32_t entry, const uint32_t v[3]) {
    if ((entry & UNIFORM_BIT) != 0u) return entry & 0xFFFFu;
    uint32_t hdr = brick_at(c, entry);
    uint32_t n = hdr & 0xFFu;
    uint32_t bits = (hdr >> 8u) & 0xFFu;
    uint32_t pal_words = (n + 1u) / 2u;
    uint32_t vb = cell_bit(v[0] & 3u, v[1] & 3u, v[2] & 3u);
    uint32_t bitpos = vb * bits;
    uint32_t w = brick_at(c, entry + 1u + pal_words + (bitpos >> 5u));
    /* WGSL shifts take the shift count modulo 32; bits <= 8 here. */
    uint32_t idx = (w >> (bitpos & 31u)) & ((1u << (bits & 31u)) - 1u);
    uint32_t pw = brick_at(c, entry + 1u + idx / 2u);
    c->st.brick_words += 3;
    return (pw >> ((idx & 1u) * 16u)) & 0xFFFFu;
}

static uint32_t leaf_attr(ctx *c, uint32_t ci, const uint32_t v[3]) {
    const vxo_chunk *ch = chunk_at(c, ci);
    uint32_t root_mask[2] = {ch->root[0], ch->root[1]};
    uint32_t l1b = cell_bit(v[0] >> 4u, v[1] >> 4u, v[2] >> 4u);
    uint32_t l1[4];
    if ((ch->root[2] & FULL_BIT) != 0u) {
        l1[0] = 0xFFFFFFFFu; l1[1] = 0xFFFFFFFFu; l1[2] = FULL_BIT; l1[3] = l1b * 64u;
    } else {
        inner_at(c, (ch->root[2] & PTR_MASK) + mask_below(root_mask, l1b), l1);
        c->st.l1_loads++;
    }
    uint32_t lfb = cell_bit((v[0] >> 2u) & 3u, (v[1] >> 2u) & 3u, (v[2] >> 2u) & 3u);
    uint32_t l1m[2] = {l1[0], l1[1]};
    c->st.attr_loads++;
    return brick_at(c, ch->attr_base + (l1[3] & 0x0FFFFFFFu) + mask_below(l1m, lfb));
}

static uint32_t block_at(ctx *c, uint32_t ci, const uint32_t v[3]) {
    const vxo_chunk *ch = chunk_at(c, ci);
    if ((ch->attr_flags & ATTR_UNIFORM) != 0u) return ch->attr_base;
    return leaf_block(c, leaf_attr(c, ci, v), v);
}

/* ---- f32 vector helpers in WGSL order ---- */
static float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static float length3(const float a[3]) { return sqrtf(dot3(a, a)); }
static float signf_(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

/* ray_dir (common.wgsl:358-368) */
static void ray_dir(const vxo_frame *f, float px, float py, float out[3]) {
    float ndc_x = ((px + f->jitter[0]) / (float)f->res[0]) * 2.0f - 1.0f;
    float ndc_y = 1.0f - ((py + f->jitter[1]) / (float)f->res[1]) * 2.0f;
    float sx = ndc_x * f->tan_half_fov * f->aspect;
    float sy = ndc_y * f->tan_half_fov;
    float d[3];
    for (int i = 0; i < 3; i++) d[i] = f->cam_fwd[i] + f->cam_right[i] * sx + f->cam_up[i] * sy;
    float len = length3(d);
    for (int i = 0; i < 3; i++) d[i] = d[i] / len;
    for (int i = 0; i < 3; i++) {
        if (fabsf(d[i]) < 1e-6f) d[i] = 1e-6f * signf_(d[i] + 1e-9f);
        out[i] = d[i];
    }
}

/* dither / ray_skips (common.wgsl:370-385) */
static float fractf_(float x) { return x - floorf(x); }
static int ray_skips(const vxo_frame *f, float fade, uint32_t px, uint32_t py) {
    if (fade >= 1.0f) return 0;
    float p0 = (float)px + 5.588238f * (float)f->frame_index;
    float p1 = (float)py;
    float d = fractf_(52.9829189f * fractf_(p0 * 0.06711056f + p1 * 0.00583715f));
    if (fade >= 0.0f) return d >= fade;
    return d < -fade;
}

/* normal_of (common.wgsl:387-398) */
static void normal_of(uint32_t id, float n[3]) {
    if (id >= NORMAL_CROSS_A) {
        float s = (id == NORMAL_CROSS_B) ? 1.0f : -1.0f;
        n[0] = 0.70710678f; n[1] = 0.0f; n[2] = s * 0.70710678f;
        return;
    }
    uint32_t axis = id >> 1u;
    float s = ((id & 1u) == 1u) ? 1.0f : -1.0f;
    n[0] = n[1] = n[2] = 0.0f;
    n[axis] = s;
}

/* hash_u32 (common.wgsl:1422-1430) */
uint32_t vxo_hash_u32(uint32_t x) {
    uint32_t h = x;
    h ^= h >> 16u;
    h *= 0x7feb352du;
    h ^= h >> 15u;
    h *= 0x846ca68bu;
    h ^= h >> 16u;
    return h;
}

/* leaf_solid (common.wgsl:514-518) */
static int leaf_solid(const vxo_spec *s, const int32_t wv[3], const uint32_t m[3]) {
    uint32_t q[3];
#if VXO_PLANT == 2
    for (int i = 0; i < 3; i++) q[i] = (uint32_t)wv[i] + m[i];
#else
    for (int i = 0; i < 3; i++) q[i] = (uint32_t)wv[i] * MICRO + m[i];
#endif
    uint32_t h = vxo_hash_u32((q[0] * 73856093u) ^ (q[1] * 19349663u) ^ (q[2] * 83492791u));
#if VXO_PLANT == 1
    return (float)(h >> 8u) * (1.0f / 16777216.0f) <= s->leaf_fill;
#else
    return (float)(h >> 8u) * (1.0f / 16777216.0f) < s->leaf_fill;
#endif
}

static int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* micro_entry (common.wgsl:520-528) */
static void micro_entry(const uint32_t uc[3], const float ro[3], const float rd[3], float t,
                        uint32_t axis, const int pos_dir[3], int snap, uint32_t out[3]) {
    int32_t m[3];
    for (int i = 0; i < 3; i++) {
        float p = (ro[i] + rd[i] * t - (float)uc[i]) * (float)MICRO;
        m[i] = clampi((int32_t)floorf(p), 0, (int32_t)MICRO - 1);
    }
#if VXO_PLANT == 3
    if (snap) m[axis] = pos_dir[axis] ? (int32_t)MICRO - 1 : 0;
#else
    if (snap) m[axis] = pos_dir[axis] ? 0 : (int32_t)MICRO - 1;
#endif
    for (int i = 0; i < 3; i++) out[i] = (uint32_t)m[i];
}

/* cutout_march (common.wgsl:537-577) */
typedef struct micro_hit { int hit; float t; uint32_t normal; uint32_t micro[3]; } micro_hit;

static micro_hit cutout_march(const vxo_spec *s, const int32_t wv[3], const uint32_t uc[3],
                              const float ro[3], const float rd[3], float t_in, uint32_t axis_in,
                              const int pos_dir[3], int snap) {
    micro_hit r;
    r.hit = 0; r.t = t_in; r.normal = 0; r.micro[0] = r.micro[1] = r.micro[2] = 0;
    float base[3] = {(float)uc[0], (float)uc[1], (float)uc[2]};
    float inv[3] = {1.0f / rd[0], 1.0f / rd[1], 1.0f / rd[2]};
    float ms = 1.0f / (float)MICRO;
    uint32_t mu[3];
    micro_entry(uc, ro, rd, t_in, axis_in, pos_dir, snap, mu);
    int32_t m[3] = {(int32_t)mu[0], (int32_t)mu[1], (int32_t)mu[2]};
    float t = t_in;
    uint32_t axis = axis_in;
    for (uint32_t i = 0; i < 3u * MICRO; i++) {
        if (m[0] < 0 || m[1] < 0 || m[2] < 0 || m[0] >= (int32_t)MICRO || m[1] >= (int32_t)MICRO ||
            m[2] >= (int32_t)MICRO)
            return r;
        uint32_t um[3] = {(uint32_t)m[0], (uint32_t)m[1], (uint32_t)m[2]};
        if (leaf_solid(s, wv, um)) {
            r.hit = 1;
            r.t = t;
            r.normal = axis * 2u + (pos_dir[axis] ? 0u : 1u);
            r.micro[0] = um[0]; r.micro[1] = um[1]; r.micro[2] = um[2];
            return r;
        }
        float tn[3];
        for (int k = 0; k < 3; k++) {
#if VXO_PLANT == 4
            float bound = (float)(pos_dir[k] ? m[k] : m[k] + 1);
#else
            float bound = (float)(pos_dir[k] ? m[k] + 1 : m[k]);
#endif
            tn[k] = (base[k] + bound * ms - ro[k]) * inv[k];
        }
        if (tn[0] <= tn[1] && tn[0] <= tn[2]) { t = tn[0]; axis = 0; }
        else if (tn[1] <= tn[2]) { t = tn[1]; axis = 1; }
        else { t = tn[2]; axis = 2; }
        m[axis] += pos_dir[axis] ? 1 : -1;
    }
    return r;
}

static int is_cutout_id(uint32_t id) { return id == LEAVES_ID || id == PINE_LEAVES_ID; }

/* voxel_occupied (common.wgsl:912-940) */
static int voxel_occupied(ctx *c, uint32_t ci, const int32_t v[3]) {
    if (v[0] < 0 || v[1] < 0 || v[2] < 0 || v[0] > 63 || v[1] > 63 || v[2] > 63) return 0;
    const vxo_chunk *ch = chunk_at(c, ci);
    uint32_t uv[3] = {(uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2]};
    uint32_t root_mask[2] = {ch->root[0], ch->root[1]};
    uint32_t l1b = cell_bit(uv[0] >> 4u, uv[1] >> 4u, uv[2] >> 4u);
    if (!mask_has(root_mask, l1b)) return 0;
    if ((ch->root[2] & FULL_BIT) != 0u) return 1;
    uint32_t l1[4];
    inner_at(c, (ch->root[2] & PTR_MASK) + mask_below(root_mask, l1b), l1);
    c->st.l1_loads++;
    if ((l1[2] & FULL_BIT) != 0u) return 1;
    uint32_t l1m[2] = {l1[0], l1[1]};
    uint32_t lfb = cell_bit((uv[0] >> 2u) & 3u, (uv[1] >> 2u) & 3u, (uv[2] >> 2u) & 3u);
    if (!mask_has(l1m, lfb)) return 0;
    uint32_t leaf[2];
    leaf_at(c, (l1[2] & PTR_MASK) + mask_below(l1m, lfb), leaf);
    c->st.leaf_loads++;
    return mask_has(leaf, cell_bit(uv[0] & 3u, uv[1] & 3u, uv[2] & 3u));
}

/* world_block_at (common.wgsl:266-277) */
static uint32_t world_block_at(ctx *c, const float p[3]) {
    const vxo_frame *f = c->f;
    int32_t cell[3];
    for (int i = 0; i < 3; i++) cell[i] = (int32_t)floorf(p[i] / 64.0f) - f->grid_min[i];
    for (int i = 0; i < 3; i++)
        if (cell[i] < 0 || cell[i] >= (int32_t)f->grid_dim[i]) return 0;
    uint32_t u[3] = {(uint32_t)cell[0], (uint32_t)cell[1], (uint32_t)cell[2]};
    uint32_t ci = grid_at(c, u[0] + u[1] * f->grid_dim[0] + u[2] * f->grid_dim[0] * f->grid_dim[1]);
    if (ci == NO_CHUNK) return 0;
    const vxo_chunk *ch = chunk_at(c, ci);
    c->st.chunk_loads++;
    int32_t v[3];
    for (int i = 0; i < 3; i++) v[i] = (int32_t)floorf((p[i] - ch->origin[i]) / ch->voxel_size);
    for (int i = 0; i < 3; i++)
        if (v[i] < 0 || v[i] >= 64) return 0;
    if (!voxel_occupied(c, ci, v)) return 0;
    uint32_t uv[3] = {(uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2]};
    return block_at(c, ci, uv);
}

static int casts_sun_shadow(uint32_t id) { return id != GLOWSTONE_ID; }

/* march_chunk (common.wgsl:579-910). Foliage tufts (cross_quad) sample the texture
 * atlas and are not restated: callers keep SPEC_FOLIAGE off, so foliage_aware is
 * always false here (common.wgsl:632-638). */
static void march_chunk(ctx *c, uint32_t ci, const float ro_w[3], const float rd[3],
                        float t_limit, int skip_water, int skip_foliage, int skip_glass,
                        int sun_ray, vxo_hit *h) {
    (void)skip_foliage;
    const vxo_spec *s = c->s;
    memset(h, 0, sizeof(*h));
    c->st.chunk_marches++;
    c->st.chunk_loads++;
    const vxo_chunk *ch = chunk_at(c, ci);
    float vs = ch->voxel_size;
    float ro[3], inv[3], bmin[3], bmax[3], t0[3], t1[3], tmin3[3], tmax3[3];
    for (int i = 0; i < 3; i++) {
        ro[i] = (ro_w[i] - ch->origin[i]) / vs;
        inv[i] = 1.0f / rd[i];
        bmin[i] = (ch->aabb_min[i] - ch->origin[i]) / vs;
        bmax[i] = (ch->aabb_max[i] - ch->origin[i]) / vs;
        t0[i] = (bmin[i] - ro[i]) * inv[i];
        t1[i] = (bmax[i] - ro[i]) * inv[i];
        tmin3[i] = fminf(t0[i], t1[i]);
        tmax3[i] = fmaxf(t0[i], t1[i]);
    }
    float t_enter = fmaxf(fmaxf(tmin3[0], tmin3[1]), tmin3[2]);
    float t_exit = fminf(fminf(tmax3[0], tmax3[1]), tmax3[2]);
    float t_end = fminf(t_exit, t_limit / vs);
    float t = fmaxf(t_enter, 0.0f);
    if (t >= t_end) return;
    uint32_t axis;
    if (tmin3[0] >= tmin3[1] && tmin3[0] >= tmin3[2]) axis = 0;
    else if (tmin3[1] >= tmin3[2]) axis = 1;
    else axis = 2;
    int pos_dir[3] = {rd[0] > 0.0f, rd[1] > 0.0f, rd[2] > 0.0f};
    int32_t cell[3];
    for (int i = 0; i < 3; i++) cell[i] = (int32_t)floorf(ro[i] + rd[i] * t);
    if (t_enter > 0.0f)
        cell[axis] = pos_dir[axis] ? (int32_t)floorf(bmin[axis]) : (int32_t)ceilf(bmax[axis]) - 1;
    for (int i = 0; i < 3; i++) cell[i] = clampi(cell[i], 0, 63);
    uint32_t root_mask[2] = {ch->root[0], ch->root[1]};
    uint32_t root_ptr = ch->root[2] & PTR_MASK;

    int water_aware = skip_water && (ch->attr_flags & ATTR_HAS_WATER) != 0u;
    if (water_aware && (ch->attr_flags & ATTR_UNIFORM) != 0u) {
        if (ch->attr_base == WATER_ID) return;
        water_aware = 0;
    }
    int attr_uniform = (ch->attr_flags & ATTR_UNIFORM) != 0u;
    int cutout_aware = 0;
    if (s->leaf_cutout) {
        cutout_aware = (ch->attr_flags & ATTR_HAS_CUTOUT) != 0u && vs == 1.0f;
        if (cutout_aware && attr_uniform && ch->attr_base != LEAVES_ID && ch->attr_base != PINE_LEAVES_ID)
            cutout_aware = 0;
    }
    int glass_aware = 0;
    if (s->glass) {
        glass_aware = skip_glass && (ch->attr_flags & ATTR_HAS_GLASS) != 0u;
        if (glass_aware && attr_uniform) {
            if (ch->attr_base == GLASS_ID) return;
            glass_aware = 0;
        }
    }
    int noncaster_aware = 0;
    if (s->emitter_gather) noncaster_aware = sun_ray && (ch->attr_flags & ATTR_HAS_EMITTER) != 0u;
    if (noncaster_aware && attr_uniform && !casts_sun_shadow(ch->attr_base)) return;
    int water_interface = s->water_repair && !skip_water && (ch->attr_flags & ATTR_HAS_WATER) != 0u;
    int attr_needed = (noncaster_aware && !attr_uniform) || (water_interface && !attr_uniform) ||
                      water_aware || glass_aware || (cutout_aware && !attr_uniform);
    uint32_t trav_mask[2] = {root_mask[0], root_mask[1]};
    if (water_aware) { trav_mask[0] = ch->dry_mask[0]; trav_mask[1] = ch->dry_mask[1]; }
    int root_full = (ch->root[2] & FULL_BIT) != 0u;
    if (root_full && !s->full_march && !noncaster_aware && !water_interface) {
        uint32_t uc[3] = {(uint32_t)cell[0], (uint32_t)cell[1], (uint32_t)cell[2]};
        h->hit = 1;
        h->t = t * vs;
        h->normal = axis * 2u + (pos_dir[axis] ? 0u : 1u);
        memcpy(h->voxel, uc, sizeof uc);
        micro_entry(uc, ro, rd, t, axis, pos_dir, t > 0.0f, h->micro);
        return;
    }
    if (trav_mask[0] == 0u && trav_mask[1] == 0u) return;
    uint32_t l1_cached = 0xFFFFFFFFu, leaf_cached = 0xFFFFFFFFu;
    uint32_t l1[4] = {0, 0, 0, 0}, leaf[2] = {0, 0}, attr = 0;
    for (;;) {
        h->iters += 1u;
        c->st.iters++;
        if (h->iters > MAX_ITERS) { c->st.capped++; break; }
        if (cell[0] < 0 || cell[1] < 0 || cell[2] < 0 || cell[0] > 63 || cell[1] > 63 || cell[2] > 63) break;
        uint32_t uc[3] = {(uint32_t)cell[0], (uint32_t)cell[1], (uint32_t)cell[2]};
        int32_t skip = 16;
        int solid = 0;
        uint32_t l1b = cell_bit(uc[0] >> 4u, uc[1] >> 4u, uc[2] >> 4u);
        if (mask_has(trav_mask, l1b)) {
            if (l1b != l1_cached) {
                l1_cached = l1b;
                if (root_full) {
                    l1[0] = 0xFFFFFFFFu; l1[1] = 0xFFFFFFFFu; l1[2] = FULL_BIT; l1[3] = (l1b * 64u) | 0xF0000000u;
                } else {
                    inner_at(c, root_ptr + mask_below(root_mask, l1b), l1);
                    c->st.l1_loads++;
                }
                leaf_cached = 0xFFFFFFFFu;
            }
            int line_skip = 0;
            if (water_aware) {
                uint32_t wl = l1[3] >> 28u;
                float ly = (float)(uc[1] & 15u);
                if (wl < 15u && ly > (float)wl) {
                    line_skip = rd[1] >= 0.0f;
                    if (!line_skip) {
                        float hh = fmaxf(fabsf(rd[0]), fabsf(rd[2]));
                        if (hh > 1e-6f && (ly - (float)wl) > 16.0f * (fabsf(rd[1]) / hh)) line_skip = 1;
                    }
                }
            }
            if (line_skip) {
                skip = 16;
            } else if ((l1[2] & FULL_BIT) != 0u) {
                solid = 1;
                if (attr_needed) {
                    uint32_t lfb = cell_bit((uc[0] >> 2u) & 3u, (uc[1] >> 2u) & 3u, (uc[2] >> 2u) & 3u);
                    if (lfb != leaf_cached) {
                        leaf_cached = lfb;
                        attr = brick_at(c, ch->attr_base + (l1[3] & 0x0FFFFFFFu) + lfb);
                        c->st.attr_loads++;
                    }
                }
            } else {
                uint32_t l1m[2] = {l1[0], l1[1]};
                uint32_t lfb = cell_bit((uc[0] >> 2u) & 3u, (uc[1] >> 2u) & 3u, (uc[2] >> 2u) & 3u);
                if (mask_has(l1m, lfb)) {
                    if (lfb != leaf_cached) {
                        leaf_cached = lfb;
                        uint32_t li = mask_below(l1m, lfb);
                        leaf_at(c, (l1[2] & PTR_MASK) + li, leaf);
                        c->st.leaf_loads++;
                        if (attr_needed) {
                            attr = brick_at(c, ch->attr_base + (l1[3] & 0x0FFFFFFFu) + li);
                            c->st.attr_loads++;
                        }
                    }
                    uint32_t vb = cell_bit(uc[0] & 3u, uc[1] & 3u, uc[2] & 3u);
                    if (mask_has(leaf, vb)) {
                        solid = 1;
                    } else {
                        uint32_t sh = vb & 42u, lo;
                        if (sh == 0u) lo = leaf[0];
                        else if (sh < 32u) lo = (leaf[0] >> sh) | (leaf[1] << (32u - sh));
                        else lo = leaf[1] >> (sh - 32u);
                        skip = ((lo & 0x00330033u) == 0u) ? 2 : 1;
                    }
                } else {
                    skip = 4;
                }
            }
        } else {
            uint32_t sh = l1b & 42u, rlo;
            if (sh == 0u) rlo = trav_mask[0];
            else if (sh < 32u) rlo = (trav_mask[0] >> sh) | (trav_mask[1] << (32u - sh));
            else rlo = trav_mask[1] >> (sh - 32u);
            skip = ((rlo & 0x00330033u) == 0u) ? 32 : 16;
        }
        if (solid && noncaster_aware) {
            uint32_t id = ch->attr_base;
            if (!attr_uniform) id = leaf_block(c, attr, uc);
            if (!casts_sun_shadow(id)) { solid = 0; skip = 1; }
        }
        if (solid && water_interface) {
            uint32_t id = ch->attr_base;
            if (!attr_uniform) id = leaf_block(c, attr, uc);
            if (id == WATER_ID) {
                uint32_t nid = axis * 2u + (pos_dir[axis] ? 0u : 1u);
                float outward[3], boundary[3], probe[3];
                normal_of(nid, outward);
                float tw = t * vs;
                float off = 0.002f * vs;
                for (int i = 0; i < 3; i++) boundary[i] = ro_w[i] + rd[i] * tw;
#if VXO_PLANT == 5
                for (int i = 0; i < 3; i++) probe[i] = boundary[i] - outward[i] * off;
#else
                for (int i = 0; i < 3; i++) probe[i] = boundary[i] + outward[i] * off;
#endif
#if VXO_PLANT != 6
                if (world_block_at(c, probe) == WATER_ID) { solid = 0; skip = 1; }
#else
                (void)probe;
#endif
            }
        }
        if (solid && water_aware) {
            if (leaf_block(c, attr, uc) == WATER_ID) {
                solid = 0;
                skip = ((attr & UNIFORM_BIT) != 0u) ? 4 : 1;
            }
        }
        if (solid && glass_aware) {
            if (leaf_block(c, attr, uc) == GLASS_ID) {
                solid = 0;
                skip = ((attr & UNIFORM_BIT) != 0u) ? 4 : 1;
            }
        }
        if (solid && cutout_aware) {
            uint32_t bid = ch->attr_base;
            if (!attr_uniform) bid = leaf_block(c, attr, uc);
            if (is_cutout_id(bid)) {
                int32_t wv[3];
                for (int i = 0; i < 3; i++) wv[i] = (int32_t)ch->origin[i] + (int32_t)uc[i];
                micro_hit q = cutout_march(s, wv, uc, ro, rd, t, axis, pos_dir, t > 0.0f);
                if (q.hit) {
                    h->hit = 1;
                    h->t = q.t * vs;
                    h->normal = q.normal;
                    memcpy(h->voxel, uc, sizeof uc);
                    memcpy(h->micro, q.micro, sizeof q.micro);
                    return;
                }
                solid = 0;
                skip = 1;
            }
        }
        if (solid) {
            h->hit = 1;
            h->t = t * vs;
            h->normal = axis * 2u + (pos_dir[axis] ? 0u : 1u);
            memcpy(h->voxel, uc, sizeof uc);
            micro_entry(uc, ro, rd, t, axis, pos_dir, t > 0.0f, h->micro);
            return;
        }
        int32_t bmin_i[3], bmax_i[3];
        float tn[3];
        for (int i = 0; i < 3; i++) {
            bmin_i[i] = cell[i] - (cell[i] & (skip - 1));
            bmax_i[i] = bmin_i[i] + skip;
            float bound = pos_dir[i] ? (float)bmax_i[i] : (float)bmin_i[i];
            tn[i] = (bound - ro[i]) * inv[i];
        }
        if (tn[0] <= tn[1] && tn[0] <= tn[2]) { t = tn[0]; axis = 0; }
        else if (tn[1] <= tn[2]) { t = tn[1]; axis = 1; }
        else { t = tn[2]; axis = 2; }
        if (t >= t_end) break;
        for (int i = 0; i < 3; i++) cell[i] = (int32_t)floorf(ro[i] + rd[i] * t);
        cell[axis] = pos_dir[axis] ? bmax_i[axis] : bmin_i[axis] - 1;
    }
}

/* shadow_ray_t (common.wgsl:1156-1195) */
static float shadow_ray_t(ctx *c, const float ro[3], const float rd[3], float max_dist) {
    const vxo_frame *f = c->f;
    float inv[3] = {1.0f / rd[0], 1.0f / rd[1], 1.0f / rd[2]};
    int pos_dir[3] = {rd[0] > 0.0f, rd[1] > 0.0f, rd[2] > 0.0f};
    int32_t cell[3];
    for (int i = 0; i < 3; i++) cell[i] = (int32_t)floorf(ro[i] / 64.0f);
    float t = 0.0f;
    uint32_t last = NO_CHUNK;
    for (uint32_t i = 0; i < 24u; i++) {
        int32_t rel[3];
        for (int k = 0; k < 3; k++) rel[k] = cell[k] - f->grid_min[k];
        for (int k = 0; k < 3; k++)
            if (rel[k] < 0 || rel[k] >= (int32_t)f->grid_dim[k]) return -1.0f;
        uint32_t u[3] = {(uint32_t)rel[0], (uint32_t)rel[1], (uint32_t)rel[2]};
        uint32_t ci = grid_at(c, u[0] + u[1] * f->grid_dim[0] + u[2] * f->grid_dim[0] * f->grid_dim[1]);
        if (ci != NO_CHUNK && ci != last) {
            last = ci;
            vxo_hit m;
            march_chunk(c, ci, ro, rd, max_dist, 1, 1, 1, 1, &m);
            if (m.hit) return m.t;
        }
        float tn[3];
        for (int k = 0; k < 3; k++) {
            float bound = (float)(pos_dir[k] ? cell[k] + 1 : cell[k]) * 64.0f;
            tn[k] = (bound - ro[k]) * inv[k];
        }
        uint32_t axis = 0;
        if (tn[0] <= tn[1] && tn[0] <= tn[2]) t = tn[0];
        else if (tn[1] <= tn[2]) { t = tn[1]; axis = 1; }
        else { t = tn[2]; axis = 2; }
        if (t >= max_dist) return -1.0f;
        cell[axis] += pos_dir[axis] ? 1 : -1;
    }
    return -1.0f;
}

/* shadow_ray (common.wgsl:1098-1148): a ray that starts under the sea level and rises
 * jumps to the sea surface first (SPEC_WATER_REPAIR), then walks the chunk grid as
 * shadow_ray_t does with the origin and range it has now; any hit blocks. shadow_ray_t's
 * hit distance is march_chunk's max(t_enter, 0) * voxel_size, never negative, so
 * "shadow_ray_t >= 0" is the walk's hit flag. */
static int shadow_ray(ctx *c, const float ro[3], const float rd[3], float max_dist) {
    const vxo_frame *f = c->f;
    float r_ro[3] = {ro[0], ro[1], ro[2]};
    float r_dist = max_dist;
    if (c->s->water_repair && rd[1] > 1e-4f && ro[1] < f->sea_level) {
        float t_surf = (f->sea_level - ro[1]) / rd[1];
        if (t_surf >= r_dist) return 0;
        for (int k = 0; k < 3; k++) r_ro[k] = ro[k] + rd[k] * t_surf;
        r_dist = r_dist - t_surf;
    }
    return shadow_ray_t(c, r_ro, rd, r_dist) >= 0.0f;
}

/* ---- secondary legs ---------------------------------------------------------------- */
#define WATER_IOR 1.333f           /* common.wgsl:1256 */
#define SURFACE_EPS 0.02f          /* common.wgsl:1258 */
#define WATER_REFRACT_NEAR 96.0f   /* resolve.wgsl:873-880, :886 */
#define WATER_REFRACT_FAR 192.0f
#define WATER_REFLECT_NEAR 512.0f
#define WATER_REFLECT_FAR 1024.0f
#define WATER_REFRACT_DIST 96.0f
#define WATER_REFLECT_DIST 768.0f
#define GLASS_TRANSMIT_DIST 192.0f
#define FLAG_WATER_REFLECT 16u     /* common.wgsl:146-147 */
#define FLAG_WATER_REFRACT 32u
#define FLAG_WATER_SHADOW_CUT 262144u  /* src/render/mod.rs:198 */
/* Sun-ray ranges of shade_hit at a leg's hit: SPEC_WATER_SHADOW_DIST for the refract leg
 * (src/render/mod.rs:325-327, :2427-2432), frame.shadow_dist for the water reflect leg,
 * GLASS_SHADOW_DIST for the glass legs (resolve.wgsl:888). */
#define WATER_SHADOW_DIST 16.0f
#define WATER_SHADOW_DIST_PRE41 48.0f
#define GLASS_SHADOW_DIST 48.0f
#define SEC_SUN_TESTED (1u << 29)
#define SEC_SUN_BLOCKED (1u << 28)

typedef struct world_hit {
    int hit;
    float t;
    uint32_t normal, voxel, ci;
} world_hit;

static float wgsl_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

/* trace_world (common.wgsl:1205-1253): grid DDA over chunk cells (at most 16) from ro,
 * marching each resident chunk with skip_glass; the first hit ends the walk. */
static world_hit trace_world(ctx *c, const float ro[3], const float rd[3], float max_dist, int skip_water) {
    const vxo_frame *f = c->f;
    world_hit out = {0, max_dist, 0u, 0u, NO_CHUNK};
    float dir[3], inv[3];
    int pos_dir[3];
    for (int k = 0; k < 3; k++) {
        int tiny = fabsf(rd[k]) < 1e-6f;
        dir[k] = tiny ? 1e-6f * wgsl_sign(rd[k] + 1e-9f) : rd[k];
        inv[k] = 1.0f / dir[k];
        pos_dir[k] = dir[k] > 0.0f;
    }
    int32_t cell[3];
    for (int k = 0; k < 3; k++) cell[k] = (int32_t)floorf(ro[k] / 64.0f);
    for (uint32_t i = 0; i < 16u; i++) {
        int32_t rel[3];
        for (int k = 0; k < 3; k++) rel[k] = cell[k] - f->grid_min[k];
        for (int k = 0; k < 3; k++)
            if (rel[k] < 0 || rel[k] >= (int32_t)f->grid_dim[k]) return out;
        uint32_t u[3] = {(uint32_t)rel[0], (uint32_t)rel[1], (uint32_t)rel[2]};
        uint32_t ci = grid_at(c, u[0] + u[1] * f->grid_dim[0] + u[2] * f->grid_dim[0] * f->grid_dim[1]);
        float tn[3];
        for (int k = 0; k < 3; k++) {
            float bound = (float)(pos_dir[k] ? cell[k] + 1 : cell[k]) * 64.0f;
            tn[k] = (bound - ro[k]) * inv[k];
        }
        float t_exit = tn[0];
        uint32_t axis = 0;
        if (tn[1] < t_exit) { t_exit = tn[1]; axis = 1; }
        if (tn[2] < t_exit) { t_exit = tn[2]; axis = 2; }
        if (ci != NO_CHUNK) {
            vxo_hit h;
            march_chunk(c, ci, ro, rd, fminf(t_exit, max_dist), skip_water, 0, 1, 0, &h);
            if (h.hit) {
                out.hit = 1;
                out.t = h.t;
                out.normal = h.normal;
                out.voxel = h.voxel[0] | (h.voxel[1] << 6) | (h.voxel[2] << 12);
                out.ci = ci;
                return out;
            }
        }
        if (t_exit >= max_dist) return out;
        cell[axis] += pos_dir[axis] ? 1 : -1;
    }
    return out;
}

/* WGSL builtins, operation order fixed: dot = (x*x' + y*y') + z*z'. */
static float dot3o(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* refract(I, N, eta) = k < 0 ? 0 : eta*I - (eta*dot(N,I) + sqrt(k))*N,
 * k = 1 - eta*eta*(1 - dot(N,I)*dot(N,I)) */
static void wgsl_refract(const float i[3], const float n[3], float eta, float out[3]) {
    float d = dot3o(n, i);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) { out[0] = out[1] = out[2] = 0.0f; return; }
    float s = eta * d + sqrtf(k);
    for (int q = 0; q < 3; q++) out[q] = eta * i[q] - s * n[q];
}

/* reflect(I, N) = I - 2*dot(N,I)*N */
static void wgsl_reflect(const float i[3], const float n[3], float out[3]) {
    float d2 = 2.0f * dot3o(n, i);
    for (int q = 0; q < 3; q++) out[q] = i[q] - d2 * n[q];
}

/* smoothstep(e0, e1, x) = t*t*(3 - 2t), t = clamp((x - e0)/(e1 - e0), 0, 1) */
static float wgsl_smoothstep(float e0, float e1, float x) {
    float t = fminf(fmaxf((x - e0) / (e1 - e0), 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

/* ---- waved water (water_ctx, resolve.wgsl:1078-1167) ------------------------------------ */
#define NORMAL_PY 3u               /* common.wgsl:295 */
#define FLAG_WAVES 512u            /* src/render/mod.rs:180-186 */
#define FLAG_WAVE_ANISO 1024u
#define FLAG_WAVE_SHOAL 2048u
#define FLAG_WAVE_FILL 4096u
#define WAVE_TAU 6.28318531f       /* resolve.wgsl:902-912 */
#define WAVE_GRAZE_MIN 1e-3f
#define SHOAL_DEPTH 6.0f
#define SHOAL_MAX 6.0f
#define WAVE_FILTER_K 1.0f

/* cos and exp: WGSL leaves their accuracy to the implementation, so the reference GPU's
 * own rounding is not pinned. This restatement and the HIP kernels use the same f32 steps:
 *   cos: k = round-to-even(x * 2/pi); r = ((x - k*A) - k*B) - k*C with A + B + C = pi/2
 *        (Cody and Waite's reduction; A and B have 12 significant bits, so k*A and k*B are
 *        exact for |k| < 4096); then the Taylor polynomial of cos (degree 8) or sin (degree
 *        9) on |r| <= pi/4, chosen and signed by k mod 4, evaluated by Horner's rule;
 *   exp: 0 for x < -86; k = round-to-even(x / ln 2), r = (x - k*L1) - k*L2 (L1 of 12 bits),
 *        the degree-7 Taylor polynomial of exp(r), times 2^k.
 * tests/test_oracle_wave_pins.py bounds both against float64. */
float vxo_wave_cos(float x) {
    const float kf = rintf(x * 0x1.45f306p-1f);                      /* 2/pi */
    const float r = ((x - kf * 0x1.922p0f) - kf * -0x1.2aep-18f) - kf * -0x1.de973ep-31f;
    const float r2 = r * r;
    float cs = 0x1.a01a02p-16f;                                       /* 1/8! */
    cs = cs * r2 - 0x1.6c16c2p-10f;                                   /* 1/6! */
    cs = cs * r2 + 0x1.555556p-5f;                                    /* 1/4! */
    cs = cs * r2 - 0.5f;
    cs = cs * r2 + 1.0f;
    float sn = 0x1.71de3ap-19f;                                       /* 1/9! */
    sn = sn * r2 - 0x1.a01a02p-13f;                                   /* 1/7! */
    sn = sn * r2 + 0x1.111112p-7f;                                    /* 1/5! */
    sn = sn * r2 - 0x1.555556p-3f;                                    /* 1/3! */
    sn = sn * r2 + 1.0f;
    sn = sn * r;
    switch ((int32_t)kf & 3) {
        case 0: return cs;
        case 1: return -sn;
        case 2: return -cs;
        default: return sn;
    }
}

/* sin with the same reduction and polynomials as vxo_wave_cos (sin = the sine branch for
 * k = 0 mod 4, then cos, -sin, -cos). */
float vxo_wave_sin(float x) {
    const float kf = rintf(x * 0x1.45f306p-1f);
    const float r = ((x - kf * 0x1.922p0f) - kf * -0x1.2aep-18f) - kf * -0x1.de973ep-31f;
    const float r2 = r * r;
    float cs = 0x1.a01a02p-16f;
    cs = cs * r2 - 0x1.6c16c2p-10f;
    cs = cs * r2 + 0x1.555556p-5f;
    cs = cs * r2 - 0.5f;
    cs = cs * r2 + 1.0f;
    float sn = 0x1.71de3ap-19f;
    sn = sn * r2 - 0x1.a01a02p-13f;
    sn = sn * r2 + 0x1.111112p-7f;
    sn = sn * r2 - 0x1.555556p-3f;
    sn = sn * r2 + 1.0f;
    sn = sn * r;
    switch ((int32_t)kf & 3) {
        case 0: return sn;
        case 1: return cs;
        case 2: return -sn;
        default: return -cs;
    }
}

float vxo_wave_exp(float x) {
    if (x < -86.0f) return 0.0f;
    const float kf = rintf(x * 0x1.715476p0f);                        /* 1/ln 2 */
    const float r = (x - kf * 0x1.62ep-1f) - kf * 0x1.0bfbe8p-15f;
    float e = 0x1.a01a02p-13f;                                        /* 1/7! */
    e = e * r + 0x1.6c16c2p-10f;
    e = e * r + 0x1.111112p-7f;
    e = e * r + 0x1.555556p-5f;
    e = e * r + 0x1.555556p-3f;
    e = e * r + 0.5f;
    e = e * r + 1.0f;
    e = e * r + 1.0f;
    return ldexpf(e, (int)kf);
}

/* normalize(v) = v / sqrt(dot(v, v)); inverseSqrt(x) = 1 / sqrt(x) */
static void wgsl_normalize(float v[3]) {
    const float len = sqrtf(dot3o(v, v));
    for (int k = 0; k < 3; k++) v[k] = v[k] / len;
}

/* wave_octave (resolve.wgsl:928-944) */
static void wave_octave(const float p[2], const float dir[2], float lambda, float slope, float phase, float px,
                        const float vh[2], float kk, int aniso, float out[3]) {
    float fade;
    if (aniso) {
        const float d = dir[0] * vh[0] + dir[1] * vh[1];
        const float ext = px * sqrtf(1.0f + kk * d * d);
        const float u = ext / lambda;
#if VXO_PLANT == 9
        fade = vxo_wave_exp(-WAVE_FILTER_K * u);
#else
        fade = vxo_wave_exp(-WAVE_FILTER_K * u * u);
#endif
    } else {
        fade = wgsl_smoothstep(2.0f, 4.0f, lambda / fmaxf(px, 1e-4f));
    }
#if VXO_PLANT == 7
    const float g = slope * fade * vxo_wave_cos(WAVE_TAU / lambda * (p[0] * dir[0] + p[1] * dir[1]) - phase);
#elif VXO_PLANT == 8
    const float g = slope * fade * vxo_wave_cos(WAVE_TAU / lambda * (p[0] * dir[1] + p[1] * dir[0]) + phase);
#else
    const float g = slope * fade * vxo_wave_cos(WAVE_TAU / lambda * (p[0] * dir[0] + p[1] * dir[1]) + phase);
#endif
    out[0] = dir[0] * g;
    out[1] = dir[1] * g;
    out[2] = 0.5f * slope * slope * (1.0f - fade * fade);
}

/* wave_field (resolve.wgsl:946-967): the octaves summed left to right */
static const float WAVE_D[4][2] = {{0.913089f, 0.407760f}, {-0.388685f, 0.921371f},
                                   {-0.980382f, -0.197108f}, {0.477328f, -0.878725f}};
static const float WAVE_D_DIV[4] = {0.0f, 2.17f, 4.71f, 10.20f};
static const float WAVE_D_AMP[4] = {0.42f, 0.27f, 0.19f, 0.12f};
static const float WAVE_D_PH[4] = {1.0000f, 1.4731f, 2.1703f, 3.1937f};
static const float WAVE_F[8][2] = {{-0.438371f, 0.898794f}, {-0.927184f, 0.374607f}, {-0.190809f, 0.981627f},
                                   {-0.669131f, 0.743145f}, {0.121869f, 0.992546f}, {0.694658f, 0.719340f},
                                   {0.970296f, 0.241922f}, {0.438371f, 0.898794f}};
static const float WAVE_F_DIV[8] = {0.0f, 1.3934f, 1.9417f, 2.7056f, 3.7700f, 5.2533f, 7.3201f, 10.2000f};
static const float WAVE_F_AMP[8] = {0.3086f, 0.2584f, 0.2163f, 0.1810f, 0.1515f, 0.1269f, 0.1062f, 0.0889f};
static const float WAVE_F_PH[8] = {1.0000f, 1.1804f, 1.3934f, 1.6449f, 1.9417f, 2.2920f, 2.7056f, 3.1937f};

void vxo_wave_field(const vxo_frame *f, const vxo_spec *s, float px, float pz, float px_world,
                    const float vh[2], float kk, int aniso, float shoal, float out[3]) {
    const float p[2] = {px, pz};
    const float l0 = f->wave_scale;
    const float a = f->wave_amp * shoal;
    const float ph = f->time * f->wave_speed * WAVE_TAU;
    const int fill = s->wave_fill && (f->flags & FLAG_WAVE_FILL) != 0u;
    const int n = fill ? 8 : 4;
    out[0] = out[1] = out[2] = 0.0f;
    for (int i = 0; i < n; i++) {
        const float *dir = fill ? WAVE_F[i] : WAVE_D[i];
        const float div = fill ? WAVE_F_DIV[i] : WAVE_D_DIV[i];
        const float lambda = i == 0 ? l0 : l0 / div;  /* the first octave is l0 itself */
        const float slope = a * (fill ? WAVE_F_AMP[i] : WAVE_D_AMP[i]);
        const float phase = ph * (fill ? WAVE_F_PH[i] : WAVE_D_PH[i]);
        float o[3];
        wave_octave(p, dir, lambda, slope, phase, px_world, vh, kk, aniso, o);
        if (i == 0) { out[0] = o[0]; out[1] = o[1]; out[2] = o[2]; }
        else { out[0] = out[0] + o[0]; out[1] = out[1] + o[1]; out[2] = out[2] + o[2]; }
    }
}

/* water_ctx's reflect-leg direction mirror_t (resolve.wgsl:1078-1167): about the face
 * normal, or on waved top faces about the wave normal at the reflect slope, blended back
 * toward the face normal by the optical roughness of the unresolved wave variance. The
 * water-look ripples (SPEC_WATER_LOOK) are not restated; callers keep them off. */
/* water_ctx's normals (resolve.wgsl:1078-1167): n the face normal; on waved top faces
 * nw the wave normal and nt the wave normal at the reflect slope (the shoal depth from a
 * short trace straight down scales the waves), gloss from the unresolved variance; both
 * blended back toward n by the optical roughness of that variance. */
typedef struct wnormals {
    float n[3], nw[3], nt[3], gloss, rough;
    int waved;
} wnormals;

static void water_normals(ctx *c, uint32_t nid, const float hit[3], const float rd[3], float t, wnormals *o) {
    const vxo_frame *f = c->f;
    const vxo_spec *s = c->s;
    normal_of(nid, o->n);
    for (int k = 0; k < 3; k++) o->nw[k] = o->nt[k] = o->n[k];
    o->gloss = 1.0f;
    float unresolved = 0.0f;
    o->waved = s->waves && (f->flags & FLAG_WAVES) != 0u && nid == NORMAL_PY;
    if (o->waved) {
        const float px_world = fmaxf(t, 1e-3f) * 2.0f * f->tan_half_fov / (float)f->res[1];
        float kk = 0.0f, vh[2] = {0.0f, 0.0f};
        const int aniso = s->wave_aniso && (f->flags & FLAG_WAVE_ANISO) != 0u;
        if (aniso) {
            const float k = 1.0f / fmaxf(fabsf(rd[1]), WAVE_GRAZE_MIN);
            kk = k * k - 1.0f;
            const float is = 1.0f / sqrtf(fmaxf(rd[0] * rd[0] + rd[2] * rd[2], 1e-12f));
            vh[0] = rd[0] * is;
            vh[1] = rd[2] * is;
        }
        float shoal = 1.0f;
        if (s->wave_shoal && (f->flags & FLAG_WAVE_SHOAL) != 0u) {
            const float dro[3] = {floorf(hit[0]) + 0.5f, hit[1] - SURFACE_EPS, floorf(hit[2]) + 0.5f};
            const float sdir[3] = {0.001f, -0.999999f, 0.001f};
            world_hit dh = trace_world(c, dro, sdir, SHOAL_MAX, 1);
            const float depth = dh.hit ? dh.t : SHOAL_MAX;
            shoal = wgsl_smoothstep(0.0f, SHOAL_DEPTH, depth);
        }
        float w[3];
        vxo_wave_field(f, s, hit[0], hit[2], px_world, vh, kk, aniso, shoal, w);
        unresolved = fmaxf(w[2], 0.0f);
        o->nw[0] = -w[0];
        o->nw[1] = 1.0f;
        o->nw[2] = -w[1];
        wgsl_normalize(o->nw);
        const float len = sqrtf(w[0] * w[0] + w[1] * w[1]);
        const float m = fminf(1.0f, f->wave_reflect_slope / fmaxf(len, 1e-6f));
        o->nt[0] = -(w[0] * m);
        o->nt[1] = 1.0f;
        o->nt[2] = -(w[1] * m);
        wgsl_normalize(o->nt);
        o->gloss = 1.0f / (1.0f + 160.0f * w[2]); /* 0.5 * GLINT_POWER */
    }
#if VXO_PLANT == 10
    const float rough = 0.0f;
    (void)unresolved;
#else
    const float rough = fminf(fmaxf(1.0f - vxo_wave_exp(-8.0f * unresolved), 0.0f), 1.0f);
#endif
    o->rough = rough;
    for (int k = 0; k < 3; k++) {
        o->nw[k] = o->nw[k] * (1.0f - rough) + o->n[k] * rough;
        o->nt[k] = o->nt[k] * (1.0f - rough) + o->n[k] * rough;
    }
    wgsl_normalize(o->nw);
    wgsl_normalize(o->nt);
}

/* water_ctx's mirror_t: reflect(rd, nt), on waved faces kept above the horizon. */
static void water_mirror_t(ctx *c, uint32_t nid, const float hit[3], const float rd[3], float t, float out[3]) {
    wnormals w;
    water_normals(c, nid, hit, rd, t, &w);
    wgsl_reflect(rd, w.nt, out);
    if (w.waved) {
        out[1] = fmaxf(out[1], 0.0f);
        wgsl_normalize(out);
    }
}

/* The sun test shade_hit makes at a leg's hit (shade_hit_compact, resolve.wgsl:587-595
 * for n and vs, :741-745; shade_hit_legacy :507-514 is the same): n = normal_of(normal),
 * turned toward the ray for cross quads; only when dot(n, sun) > 0, daylight > 0.01 and
 * FLAG_SHADOWS does shadow_ray run, from hit + n * (0.02 * vs). A glass leg that ends on
 * water is shaded by shade_water instead (resolve.wgsl:1284-1289, :1305-1309): no sun test
 * at that hit (the transmit leg's water legs are records 2 and 3, tested at their own). */
static uint32_t leg_sun(ctx *c, const float ro[3], const float rd[3], const world_hit *w, uint32_t material,
                        int water_leg, float shadow_dist) {
    const vxo_frame *f = c->f;
    if (!w->hit || (!water_leg && material == WATER_ID)) return 0u;
    float n[3], hit[3], o[3];
    normal_of(w->normal, n);
    if (w->normal >= NORMAL_CROSS_A && dot3o(n, rd) > 0.0f)
        for (int k = 0; k < 3; k++) n[k] = -n[k];
    float ndl = dot3o(n, f->sun_dir);
    if (!(ndl > 0.0f && f->daylight > 0.01f) || (f->flags & FLAG_SHADOWS) == 0u) return 0u;
    float vs = chunk_at(c, w->ci)->voxel_size;
    for (int k = 0; k < 3; k++) hit[k] = ro[k] + rd[k] * w->t;
    for (int k = 0; k < 3; k++) o[k] = hit[k] + n[k] * (0.02f * vs);
    c->st.rays_secondary_sun++;
    return SEC_SUN_TESTED | (shadow_ray(c, o, f->sun_dir, shadow_dist) ? SEC_SUN_BLOCKED : 0u);
}

static void store_leg(ctx *c, const float ro[3], const float rd[3], const world_hit *w, int water_leg,
                      float shadow_dist, vxo_sec *o) {
    o->t = w->t;
    o->voxel = w->hit ? w->voxel : 0u;
    o->material = 0u;
    if (w->hit) {
        uint32_t v[3] = {w->voxel & 63u, (w->voxel >> 6) & 63u, (w->voxel >> 12) & 63u};
        o->material = block_at(c, w->ci, v);
    }
    o->info = (w->hit ? (w->ci & 0x1FFFu) : 0x1FFFu) | ((w->hit ? w->normal : 0u) << 13) | (1u << 30) |
              (w->hit ? (1u << 31) : 0u) | leg_sun(c, ro, rd, w, o->material, water_leg, shadow_dist);
    c->st.rays_secondary++;
}

static float hit_t(ctx *c, uint32_t ci, const uint32_t v[3], const uint32_t micro[3], uint32_t nid,
                   const float rd[3]);

/* shade_water's two legs from a water hit (resolve.wgsl:1209-1220): the refract leg about
 * the face normal (water_refract_leg, :1029-1042) into r0, the reflect leg along water_ctx's
 * mirror_t (water_reflect_leg, :1044-1058) into r1. */
static void water_legs(ctx *c, uint32_t nid, const float hit[3], const float rd[3], float t, vxo_sec *r0,
                       vxo_sec *r1) {
    const vxo_frame *f = c->f;
    float n[3];
    normal_of(nid, n);
    float refract_detail = 1.0f - wgsl_smoothstep(WATER_REFRACT_NEAR, WATER_REFRACT_FAR, t);
    float reflect_detail = 1.0f - wgsl_smoothstep(WATER_REFLECT_NEAR, WATER_REFLECT_FAR, t);
    if ((f->flags & FLAG_WATER_REFRACT) != 0u && refract_detail > 0.0f) {
        float rdir[3], ro2[3];
        wgsl_refract(rd, n, 1.0f / WATER_IOR, rdir);
        for (int k = 0; k < 3; k++) ro2[k] = hit[k] + rdir[k] * SURFACE_EPS;
        world_hit w = trace_world(c, ro2, rdir, WATER_REFRACT_DIST, 1);
        float sd = (f->flags & FLAG_WATER_SHADOW_CUT) ? WATER_SHADOW_DIST : WATER_SHADOW_DIST_PRE41;
        store_leg(c, ro2, rdir, &w, 1, sd, r0);
    }
    if ((f->flags & FLAG_WATER_REFLECT) != 0u && reflect_detail > 0.0f) {
        float m[3], ro3[3];
        water_mirror_t(c, nid, hit, rd, t, m);
        for (int k = 0; k < 3; k++) ro3[k] = hit[k] + n[k] * SURFACE_EPS;
        world_hit w = trace_world(c, ro3, m, WATER_REFLECT_DIST, 1);
        store_leg(c, ro3, m, &w, 1, f->shadow_dist, r1);
    }
}

/* The legs of one pixel: shade_water (resolve.wgsl:1209-1220) and shade_glass (:1261-1310),
 * whose transmit leg, when it ends on water, is shaded by shade_water at that point with
 * t + the leg's distance (:1284-1289): its two legs go to out[2] and out[3]. The glass
 * reflect leg (SPEC_GLASS_REFLECT) ending on water is not continued. */
static void secondary_pixel(ctx *c, const uint64_t *vis, uint32_t px, uint32_t py, vxo_sec out[4]) {
    const vxo_frame *f = c->f;
    memset(out, 0, 4 * sizeof(vxo_sec));
    uint64_t key = vis[(size_t)py * f->res[0] + px];
    if (key == 0) return;
    uint32_t ci = (uint32_t)(key >> 24u) & 0x1FFFu;
    uint32_t low = (uint32_t)(key & 0xFFFFFFu);
    uint32_t v[3] = {(low >> 16u) & VOXEL_MASK, (low >> 8u) & VOXEL_MASK, low & VOXEL_MASK};
    uint32_t micro[3] = {(low >> 22u) & 3u, (low >> 14u) & 3u, (low >> 6u) & 3u};
    uint32_t nid = (uint32_t)(key >> 37u) & 7u;
    uint32_t id = block_at(c, ci, v);
    int glass = c->s->glass && id == GLASS_ID;
    if (id != WATER_ID && !glass) return;
    float rd[3], n[3], hit[3];
    ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd);
    float t = hit_t(c, ci, v, micro, nid, rd);
    for (int k = 0; k < 3; k++) hit[k] = f->cam_pos[k] + rd[k] * t;
    normal_of(nid, n);
    if (id == WATER_ID) {
        water_legs(c, nid, hit, rd, t, &out[0], &out[1]);
    } else {
        if (dot3o(n, rd) > 0.0f)
            for (int k = 0; k < 3; k++) n[k] = -n[k];
        float dist = fmaxf(GLASS_TRANSMIT_DIST, f->far_);
        float ro2[3];
        for (int k = 0; k < 3; k++) ro2[k] = hit[k] + rd[k] * SURFACE_EPS;
        world_hit w = trace_world(c, ro2, rd, dist, 0);
        store_leg(c, ro2, rd, &w, 0, GLASS_SHADOW_DIST, &out[0]);
        if (w.hit && out[0].material == WATER_ID) {
            float p2[3];
            for (int k = 0; k < 3; k++) p2[k] = ro2[k] + rd[k] * w.t;
            water_legs(c, w.normal, p2, rd, t + w.t, &out[2], &out[3]);
        }
        if (c->s->glass_reflect) {
            float m[3], o[3];
            for (int k = 0; k < 3; k++) o[k] = hit[k] + n[k] * SURFACE_EPS;
            wgsl_reflect(rd, n, m);
            world_hit w2 = trace_world(c, o, m, dist, 0);
            store_leg(c, o, m, &w2, 0, GLASS_SHADOW_DIST, &out[1]);
        }
    }
}

/* hit_t (common.wgsl:1300-1318) */
static float hit_t(ctx *c, uint32_t ci, const uint32_t v[3], const uint32_t micro[3], uint32_t nid,
                   const float rd[3]) {
    const vxo_frame *f = c->f;
    const vxo_chunk *ch = chunk_at(c, ci);
    float box_min[3];
    for (int i = 0; i < 3; i++) box_min[i] = ch->origin[i] + (float)v[i] * ch->voxel_size;
    if (nid >= NORMAL_CROSS_A) {
        if (nid == NORMAL_CROSS_A)
            return ((box_min[0] - box_min[2]) - (f->cam_pos[0] - f->cam_pos[2])) / (rd[0] - rd[2]);
        return ((box_min[0] + box_min[2] + ch->voxel_size) - (f->cam_pos[0] + f->cam_pos[2])) /
               (rd[0] + rd[2]);
    }
    uint32_t axis = nid >> 1u;
    float ms = ch->voxel_size / (float)MICRO;
    float plane = box_min[axis] + (float)(micro[axis] + (((nid & 1u) == 1u) ? 1u : 0u)) * ms;
    return (plane - f->cam_pos[axis]) / rd[axis];
}

/* depth_key / make_key (common.wgsl:1320-1330) */
uint32_t vxo_depth_key(const vxo_frame *f, float t) {
    float x = 1.0f - t / f->far_;
    x = fminf(fmaxf(x, 0.0f), 1.0f);
    return (uint32_t)(x * 16777215.0f);
}

uint64_t vxo_make_key(const vxo_frame *f, float t, uint32_t normal, uint32_t ci, const uint32_t v[3],
                      const uint32_t micro[3]) {
    uint32_t d = vxo_depth_key(f, t);
    uint32_t b[3];
    for (int i = 0; i < 3; i++) b[i] = v[i] | (micro[i] << MICRO_SHIFT);
    return ((uint64_t)d << 40u) | ((uint64_t)normal << 37u) | ((uint64_t)ci << 24u) |
           (uint64_t)((b[0] << 16u) | (b[1] << 8u) | b[2]);
}

/* tile_select helpers (tile_select.wgsl:1-9) */
static int box_outside(const float bmin[3], const float bmax[3], const float n[3], const float cam[3]) {
    float q[3];
    for (int i = 0; i < 3; i++) q[i] = (n[i] > 0.0f ? bmax[i] : bmin[i]) - cam[i];
    return dot3(n, q) < 0.0f;
}

static float box_distance(const float p[3], const float bmin[3], const float bmax[3]) {
    float q[3];
    for (int i = 0; i < 3; i++) q[i] = fmaxf(fmaxf(bmin[i] - p[i], p[i] - bmax[i]), 0.0f);
    return length3(q);
}

static void cross3(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* ---- per-tile frame pipeline ---- */
typedef struct job {
    const vxo_pools *p;
    const vxo_chunk *chunks;
    uint32_t n_chunks;
    const uint32_t *grid;
    const vxo_frame *f;
    const vxo_spec *s;
    uint32_t *hiz;
    uint64_t *vis;
    uint16_t *shadow;
    uint32_t *heat;        /* FLAG_HEATMAP frames: marched iterations per pixel, or NULL */
    int do_march;
    int next;              /* atomic work counter (rows of tiles or rows of pixels) */
    int nthreads;
    vxo_stats *per_thread;
} job;

static void ctx_init(ctx *c, const job *j) {
    memset(c, 0, sizeof(*c));
    c->p = j->p; c->chunks = j->chunks; c->n_chunks = j->n_chunks;
    c->grid = j->grid; c->f = j->f; c->s = j->s;
}

/* March every pixel of one tile against chunk ci (march.wgsl:1-55, Snell bend off).
 * tileheat, when non-NULL, receives each pixel's h.iters (the FLAG_HEATMAP atomicAdd
 * into dbg, march.wgsl:51-53). */
/* The Snell bend (march.wgsl:23-45): an underwater camera's ray that misses chunk ci going
 * up is refracted at the sea surface (from water: eta = WATER_IOR about (0, -1, 0)) and
 * marched again through ci from SURFACE_EPS past the surface, skipping water, within
 * far - t_surf; a hit replaces the miss with t = t_surf + its distance, iterations added. */
static void snell_bend(ctx *c, uint32_t ci, const float rd[3], vxo_hit *h) {
    const vxo_frame *f = c->f;
    const float depth = f->sea_level - f->cam_pos[1];
    if (!(depth > 0.0f)) return;
    const float t_surf = depth / rd[1];
    if (!(t_surf < f->far_)) return;
    const float n[3] = {0.0f, -1.0f, 0.0f};
    float rd2[3], ro2[3];
#if VXO_PLANT == 16
    wgsl_refract(rd, n, 1.0f / WATER_IOR, rd2);
#else
    wgsl_refract(rd, n, WATER_IOR, rd2);
#endif
    if (!(dot3o(rd2, rd2) > 0.0f)) return;
    for (int k = 0; k < 3; k++) ro2[k] = (f->cam_pos[k] + rd[k] * t_surf) + rd2[k] * SURFACE_EPS;
    vxo_hit h2;
    march_chunk(c, ci, ro2, rd2, f->far_ - t_surf, 1, 0, 0, 0, &h2);
    if (h2.hit) {
        const uint32_t straight = h->iters;
        *h = h2;
#if VXO_PLANT == 17
        h->t = h2.t;
#else
        h->t = t_surf + h2.t;
#endif
        h->iters = straight + h2.iters;
    }
}

static void march_pair(ctx *c, uint32_t tile, uint32_t ci, uint64_t *tilevis, uint32_t *tileheat) {
    const vxo_frame *f = c->f;
    uint32_t tx = tile % f->tiles[0], ty = tile / f->tiles[0];
    int underwater = (f->flags & FLAG_UNDERWATER) != 0u;
    for (uint32_t ly = 0; ly < 8; ly++) {
        for (uint32_t lx = 0; lx < 8; lx++) {
            uint32_t px = tx * 8u + lx, py = ty * 8u + ly;
            if (px >= f->res[0] || py >= f->res[1]) continue;
            if (ray_skips(f, chunk_at(c, ci)->fade, px, py)) continue;
            float rd[3];
            ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd);
            vxo_hit h;
            c->st.rays_primary++;
            march_chunk(c, ci, f->cam_pos, rd, f->far_, underwater, 0, 0, 0, &h);
            if (c->s->snell_bend && underwater && !h.hit && rd[1] > 0.0f) snell_bend(c, ci, rd, &h);
            if (h.hit) {
                uint64_t k = vxo_make_key(f, h.t, h.normal, ci, h.voxel, h.micro);
                if (k > tilevis[ly * 8 + lx]) tilevis[ly * 8 + lx] = k;
            }
            if (tileheat) tileheat[ly * 8 + lx] += h.iters;
        }
    }
}

/* tile_select (tile_select.wgsl:11-92) for one tile; selected[] / deferred[] in ci order. */
static void tile_select_one(ctx *c, uint32_t gx, uint32_t gy, uint32_t prev_hiz, uint32_t *sel,
                            uint32_t *nsel, uint32_t *def, uint32_t *ndef) {
    const vxo_frame *f = c->f;
    float x0 = (float)(gx * 8u), y0 = (float)(gy * 8u);
    float x1 = x0 + 8.0f, y1 = y0 + 8.0f;
    float d00[3], d10[3], d11[3], d01[3], center[3];
    ray_dir(f, x0, y0, d00);
    ray_dir(f, x1, y0, d10);
    ray_dir(f, x1, y1, d11);
    ray_dir(f, x0, y1, d01);
    ray_dir(f, (x0 + x1) * 0.5f, (y0 + y1) * 0.5f, center);
    float n[4][3];
    cross3(d00, d10, n[0]);
    cross3(d10, d11, n[1]);
    cross3(d11, d01, n[2]);
    cross3(d01, d00, n[3]);
    for (int k = 0; k < 4; k++)
        if (!(dot3(n[k], center) >= 0.0f))
            for (int i = 0; i < 3; i++) n[k][i] = -n[k][i];
    const float *cam = f->cam_pos;
    int use_hiz = (f->flags & FLAG_HIZ) != 0u;
    float reach = f->far_;
    if ((f->flags & (FLAG_UNDERWATER | FLAG_WATER_FAR)) == (FLAG_UNDERWATER | FLAG_WATER_FAR)) {
        float up = fmaxf(fmaxf(d00[1], d10[1]), fmaxf(d11[1], d01[1]));
        float submerged = fmaxf(f->sea_level - cam[1], 0.0f);
        float escape = 1e30f;
        if (up > 0.0f) escape = submerged / up;
        if (escape >= WATER_FAR_DIST) {
            reach = fminf(f->far_, WATER_FAR_DIST);
            float dep_far = submerged - up * WATER_FAR_DIST;
            if ((f->flags & FLAG_WATER_DARK) != 0u && fminf(submerged, dep_far) >= WATER_DARK_DEPTH)
                reach = fminf(reach, WATER_DARK_DIST);
        }
    }
    *nsel = 0; *ndef = 0;
    for (uint32_t ci = 0; ci < f->chunk_count; ci++) {
        const vxo_chunk *ch = chunk_at(c, ci);
        c->st.chunk_loads++;
        if (ci == f->camera_chunk) { sel[(*nsel)++] = ci; continue; }
        if (box_outside(ch->aabb_min, ch->aabb_max, f->cam_fwd, cam)) continue;
        int out = 0;
        for (int k = 0; k < 4 && !out; k++) out = box_outside(ch->aabb_min, ch->aabb_max, n[k], cam);
        if (out) continue;
        float dmin = box_distance(cam, ch->aabb_min, ch->aabb_max);
        if (dmin >= reach) continue;
        if (use_hiz && vxo_depth_key(f, dmin) < prev_hiz) { def[(*ndef)++] = ci; continue; }
        sel[(*nsel)++] = ci;
    }
}

static void *frame_worker(void *arg) {
    job *j = (job *)arg;
    int tid = __atomic_fetch_add(&j->nthreads, 1, __ATOMIC_SEQ_CST);
    ctx c;
    ctx_init(&c, j);
    const vxo_frame *f = j->f;
    uint32_t *sel = (uint32_t *)malloc(sizeof(uint32_t) * (f->chunk_count + 1));
    uint32_t *def = (uint32_t *)malloc(sizeof(uint32_t) * (f->chunk_count + 1));
    uint64_t tv[64];
    uint32_t th[64];
    uint32_t *heat = (j->heat && (f->flags & FLAG_HEATMAP) != 0u) ? th : NULL;
    for (;;) {
        int row = __atomic_fetch_add(&j->next, 1, __ATOMIC_SEQ_CST);
        if ((uint32_t)row >= f->tiles[1]) break;
        uint32_t gy = (uint32_t)row;
        for (uint32_t gx = 0; gx < f->tiles[0]; gx++) {
            uint32_t tile = gy * f->tiles[0] + gx;
            int use_hiz = (f->flags & FLAG_HIZ) != 0u;
            uint32_t prev = use_hiz ? j->hiz[tile] : 0u;
            uint32_t nsel, ndef;
            tile_select_one(&c, gx, gy, prev, sel, &nsel, def, &ndef);
            c.st.pairs_select += nsel;
            c.st.pairs_deferred += ndef;
            memset(tv, 0, sizeof tv);
            memset(th, 0, sizeof th);
            for (uint32_t k = 0; k < nsel; k++) march_pair(&c, tile, sel[k], tv, heat);
            if (use_hiz) {
                /* build_hiz (march.wgsl:57-76): min of the 24-bit depth, 0 off-frame. */
                uint32_t mn = 0xFFFFFFFFu;
                for (uint32_t ly = 0; ly < 8; ly++)
                    for (uint32_t lx = 0; lx < 8; lx++) {
                        uint32_t px = gx * 8u + lx, py = gy * 8u + ly;
                        uint32_t d = 0;
                        if (px < f->res[0] && py < f->res[1]) d = (uint32_t)(tv[ly * 8 + lx] >> 40u);
                        if (d < mn) mn = d;
                    }
                j->hiz[tile] = mn;
                /* recover_select (tile_select.wgsl:94-110). */
                for (uint32_t k = 0; k < ndef; k++) {
                    const vxo_chunk *ch = chunk_at(&c, def[k]);
                    float dmin = box_distance(f->cam_pos, ch->aabb_min, ch->aabb_max);
                    if (vxo_depth_key(f, dmin) >= mn) {
                        c.st.pairs_recover++;
                        march_pair(&c, tile, def[k], tv, heat);
                    }
                }
            } else {
                j->hiz[tile] = 0u;
            }
            for (uint32_t ly = 0; ly < 8; ly++)
                for (uint32_t lx = 0; lx < 8; lx++) {
                    uint32_t px = gx * 8u + lx, py = gy * 8u + ly;
                    if (px < f->res[0] && py < f->res[1]) {
                        j->vis[(size_t)py * f->res[0] + px] = tv[ly * 8 + lx];
                        if (j->heat) j->heat[(size_t)py * f->res[0] + px] = th[ly * 8 + lx];
                    }
                }
        }
    }
    free(sel);
    free(def);
    j->per_thread[tid] = c.st;
    return NULL;
}

/* primary_shadow + shadow_guide + shadow_active (shadow.wgsl:8-48) */
static uint16_t shadow_pixel(ctx *c, const uint64_t *vis, uint32_t px, uint32_t py) {
    const vxo_frame *f = c->f;
    float distance = 0.0f;
    int active = (f->flags & FLAG_SHADOWS) != 0u && (f->flags & FLAG_HEATMAP) == 0u && f->daylight > 0.01f;
    if (active) {
        uint64_t key = vis[(size_t)py * f->res[0] + px];
        if (key != 0) {
            uint32_t ci = (uint32_t)(key >> 24u) & 0x1FFFu;
            uint32_t low = (uint32_t)(key & 0xFFFFFFu);
            uint32_t v[3] = {(low >> 16u) & VOXEL_MASK, (low >> 8u) & VOXEL_MASK, low & VOXEL_MASK};
            uint32_t id = block_at(c, ci, v);
            if (!(id == WATER_ID || (c->s->glass && id == GLASS_ID))) {
                uint32_t nid = (uint32_t)(key >> 37u) & 7u;
                uint32_t micro[3] = {(low >> 22u) & 3u, (low >> 14u) & 3u, (low >> 6u) & 3u};
                float rd[3], n[3];
                ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd);
                normal_of(nid, n);
                if (nid >= NORMAL_CROSS_A && dot3(n, rd) > 0.0f)
                    for (int i = 0; i < 3; i++) n[i] = -n[i];
                if (dot3(n, f->sun_dir) > 0.0f) {
                    float t = hit_t(c, ci, v, micro, nid, rd);
                    if (t >= 0.0f) {
                        float scale = chunk_at(c, ci)->voxel_size;
                        float hit[3], o[3];
                        float e = 0.02f * scale;
                        for (int i = 0; i < 3; i++) hit[i] = f->cam_pos[i] + rd[i] * t;
                        for (int i = 0; i < 3; i++) o[i] = hit[i] + n[i] * e;
                        c->st.rays_shadow++;
                        float blocker = shadow_ray_t(c, o, f->sun_dir, f->shadow_dist);
                        if (blocker >= 0.0f) distance = fminf(fmaxf(blocker, 0.001f), 65504.0f);
                    }
                }
            }
        }
    }
    return vxo_f32_to_f16(distance);
}

typedef struct sjob {
    job base;
    const uint64_t *vis_in;
} sjob;

static void *shadow_worker(void *arg) {
    sjob *sj = (sjob *)arg;
    job *j = &sj->base;
    int tid = __atomic_fetch_add(&j->nthreads, 1, __ATOMIC_SEQ_CST);
    ctx c;
    ctx_init(&c, j);
    const vxo_frame *f = j->f;
    for (;;) {
        int row = __atomic_fetch_add(&j->next, 1, __ATOMIC_SEQ_CST);
        if ((uint32_t)row >= f->res[1]) break;
        for (uint32_t px = 0; px < f->res[0]; px++)
            j->shadow[(size_t)row * f->res[0] + px] = shadow_pixel(&c, sj->vis_in, px, (uint32_t)row);
    }
    j->per_thread[tid] = c.st;
    return NULL;
}

int vxo_hardware_threads(void) {
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

static void sum_stats(vxo_stats *dst, const vxo_stats *src, int n) {
    for (int t = 0; t < n; t++) {
        const uint64_t *a = (const uint64_t *)&src[t];
        uint64_t *d = (uint64_t *)dst;
        for (size_t k = 0; k < sizeof(vxo_stats) / 8; k++) d[k] += a[k];
    }
}

static void finish_bytes(vxo_stats *st) {
    st->node_bytes = 96ull * st->chunk_loads + 16ull * st->l1_loads + 8ull * st->leaf_loads +
                     4ull * (st->attr_loads + st->brick_words + st->grid_loads);
}

static void run_threads(void *(*fn)(void *), void *arg, int n) {
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n);
    for (int i = 0; i < n; i++) pthread_create(&th[i], NULL, fn, arg);
    for (int i = 0; i < n; i++) pthread_join(th[i], NULL);
    free(th);
}

int vxo_primary_shadow(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const uint32_t *grid,
                       const vxo_frame *f, const vxo_spec *s, const uint64_t *vis, uint16_t *shadow,
                       vxo_stats *stats, int threads) {
    if (!p || !chunks || !grid || !f || !s || !vis || !shadow) return -1;
    int n = threads > 0 ? threads : vxo_hardware_threads();
    sjob sj;
    memset(&sj, 0, sizeof sj);
    sj.base.p = p; sj.base.chunks = chunks; sj.base.n_chunks = n_chunks; sj.base.grid = grid;
    sj.base.f = f; sj.base.s = s; sj.base.shadow = shadow;
    sj.base.per_thread = (vxo_stats *)calloc((size_t)n, sizeof(vxo_stats));
    sj.vis_in = vis;
    run_threads(shadow_worker, &sj, n);
    if (stats) {
        sum_stats(stats, sj.base.per_thread, n);
        finish_bytes(stats);
    }
    free(sj.base.per_thread);
    return 0;
}

typedef struct secjob {
    job base;
    const uint64_t *vis_in;
    vxo_sec *out;
} secjob;

static void *secondary_worker(void *arg) {
    secjob *sj = (secjob *)arg;
    job *j = &sj->base;
    int tid = __atomic_fetch_add(&j->nthreads, 1, __ATOMIC_SEQ_CST);
    ctx c;
    ctx_init(&c, j);
    const vxo_frame *f = j->f;
    for (;;) {
        int row = __atomic_fetch_add(&j->next, 1, __ATOMIC_SEQ_CST);
        if ((uint32_t)row >= f->res[1]) break;
        for (uint32_t px = 0; px < f->res[0]; px++)
            secondary_pixel(&c, sj->vis_in, px, (uint32_t)row, &sj->out[4 * ((size_t)row * f->res[0] + px)]);
    }
    j->per_thread[tid] = c.st;
    return NULL;
}

int vxo_secondary(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const uint32_t *grid,
                  const vxo_frame *f, const vxo_spec *s, const uint64_t *vis, vxo_sec *out,
                  vxo_stats *stats, int threads) {
    if (!p || !chunks || !grid || !f || !s || !vis || !out) return -1;
    int n = threads > 0 ? threads : vxo_hardware_threads();
    secjob sj;
    memset(&sj, 0, sizeof sj);
    sj.base.p = p; sj.base.chunks = chunks; sj.base.n_chunks = n_chunks; sj.base.grid = grid;
    sj.base.f = f; sj.base.s = s;
    sj.base.per_thread = (vxo_stats *)calloc((size_t)n, sizeof(vxo_stats));
    sj.vis_in = vis;
    sj.out = out;
    run_threads(secondary_worker, &sj, n);
    if (stats) {
        memset(stats, 0, sizeof *stats);
        sum_stats(stats, sj.base.per_thread, n);
        finish_bytes(stats);
    }
    free(sj.base.per_thread);
    return 0;
}

int vxo_render(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const uint32_t *grid,
               const vxo_frame *f, const vxo_spec *s, uint32_t *hiz, uint64_t *vis, uint16_t *shadow,
               vxo_stats *stats, int threads) {
    return vxo_render_heat(p, chunks, n_chunks, grid, f, s, hiz, vis, shadow, NULL, stats, threads);
}

int vxo_render_heat(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const uint32_t *grid,
                    const vxo_frame *f, const vxo_spec *s, uint32_t *hiz, uint64_t *vis, uint16_t *shadow,
                    uint32_t *heat, vxo_stats *stats, int threads) {
    if (!p || !chunks || !grid || !f || !s || !hiz || !vis) return -1;
    if (f->chunk_count > n_chunks) return -1;
    int n = threads > 0 ? threads : vxo_hardware_threads();
    if (stats) memset(stats, 0, sizeof(*stats));
    job j;
    memset(&j, 0, sizeof j);
    j.p = p; j.chunks = chunks; j.n_chunks = n_chunks; j.grid = grid; j.f = f; j.s = s;
    j.hiz = hiz; j.vis = vis; j.heat = heat;
    j.per_thread = (vxo_stats *)calloc((size_t)n, sizeof(vxo_stats));
    run_threads(frame_worker, &j, n);
    if (stats) sum_stats(stats, j.per_thread, n);
    free(j.per_thread);
    if (shadow) {
        vxo_stats sst;
        memset(&sst, 0, sizeof sst);
        vxo_primary_shadow(p, chunks, n_chunks, grid, f, s, vis, shadow, &sst, threads);
        if (stats) sum_stats(stats, &sst, 1);
    }
    if (stats) finish_bytes(stats);
    return 0;
}

int vxo_decode(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const vxo_frame *f,
               const uint64_t *vis, vxo_decoded *out, int threads) {
    (void)threads;
    if (!p || !chunks || !f || !vis || !out) return -1;
    job j;
    memset(&j, 0, sizeof j);
    j.p = p; j.chunks = chunks; j.n_chunks = n_chunks; j.f = f;
    ctx c;
    ctx_init(&c, &j);
    for (uint32_t py = 0; py < f->res[1]; py++)
        for (uint32_t px = 0; px < f->res[0]; px++) {
            size_t i = (size_t)py * f->res[0] + px;
            uint64_t key = vis[i];
            vxo_decoded *o = &out[i];
            memset(o, 0, sizeof *o);
            if (key == 0) { o->t = -1.0f; continue; }
            uint32_t ci = (uint32_t)(key >> 24u) & 0x1FFFu;
            uint32_t low = (uint32_t)(key & 0xFFFFFFu);
            uint32_t v[3] = {(low >> 16u) & VOXEL_MASK, (low >> 8u) & VOXEL_MASK, low & VOXEL_MASK};
            uint32_t micro[3] = {(low >> 22u) & 3u, (low >> 14u) & 3u, (low >> 6u) & 3u};
            uint32_t nid = (uint32_t)(key >> 37u) & 7u;
            const vxo_chunk *ch = chunk_at(&c, ci);
            float rd[3];
            ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd);
            for (int k = 0; k < 3; k++) o->voxel[k] = (int32_t)ch->origin[k] + (int32_t)v[k] * (int32_t)ch->voxel_size;
            o->face = nid;
            o->material = block_at(&c, ci, v);
            o->ci = ci;
            o->t = hit_t(&c, ci, v, micro, nid, rd);
        }
    return 0;
}

/* ---- exported single-function restatements ---- */
void vxo_ray_dir(const vxo_frame *f, float px, float py, float out[3]) { ray_dir(f, px, py, out); }

uint32_t vxo_block_at(const vxo_pools *p, const vxo_chunk *chunks, uint32_t ci, const uint32_t v[3]) {
    job j;
    memset(&j, 0, sizeof j);
    j.p = p; j.chunks = chunks; j.n_chunks = ci + 1;
    ctx c;
    ctx_init(&c, &j);
    return block_at(&c, ci, v);
}

void vxo_march_chunk(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const uint32_t *grid,
                     const vxo_frame *f, const vxo_spec *s, uint32_t ci, const float ro_w[3], const float rd[3],
                     float t_limit, int skip_water, int skip_foliage, int skip_glass, int sun_ray, vxo_hit *out) {
    job j;
    memset(&j, 0, sizeof j);
    j.p = p; j.chunks = chunks; j.n_chunks = n_chunks; j.grid = grid; j.f = f; j.s = s;
    ctx c;
    ctx_init(&c, &j);
    march_chunk(&c, ci, ro_w, rd, t_limit, skip_water, skip_foliage, skip_glass, sun_ray, out);
}

void vxo_water_reflect_dir(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const uint32_t *grid,
                           const vxo_frame *f, const vxo_spec *s, uint32_t nid, const float hit[3],
                           const float rd[3], float t, float out[3]) {
    job j;
    memset(&j, 0, sizeof j);
    j.p = p; j.chunks = chunks; j.n_chunks = n_chunks; j.grid = grid; j.f = f; j.s = s;
    ctx c;
    ctx_init(&c, &j);
    water_mirror_t(&c, nid, hit, rd, t, out);
}

float vxo_shadow_ray_t(const vxo_pools *p, const vxo_chunk *chunks, uint32_t n_chunks, const uint32_t *grid,
                       const vxo_frame *f, const vxo_spec *s, const float ro[3], const float rd[3],
                       float max_dist) {
    job j;
    memset(&j, 0, sizeof j);
    j.p = p; j.chunks = chunks; j.n_chunks = n_chunks; j.grid = grid; j.f = f; j.s = s;
    ctx c;
    ctx_init(&c, &j);
    return shadow_ray_t(&c, ro, rd, max_dist);
}

/* f32 -> f16, IEEE 754 round-to-nearest-even, written from the format definition. */
uint16_t vxo_f32_to_f16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    uint32_t sign = (u >> 16) & 0x8000u;
    uint32_t exp = (u >> 23) & 0xFFu;
    uint32_t man = u & 0x7FFFFFu;
    if (exp == 0xFFu) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u : 0u));
    int32_t e = (int32_t)exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        uint32_t m = man | 0x800000u;
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t half = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t mid = 1u << (shift - 1u);
        if (rem > mid || (rem == mid && (half & 1u))) half++;
        return (uint16_t)(sign | half);
    }
    uint32_t half = ((uint32_t)e << 10) | (man >> 13);
    uint32_t rem = man & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (half & 1u))) half++;
    return (uint16_t)(sign | half);
}

#include "vxresolve.inc"
