// resolve.hip -- resolve lane 0 (resolve.wgsl:1378-1466) over the last vx_render frame:
// the heatmap frame, sky pixels, and opaque pixels shaded from the key with the sun
// visibility of the shadow passes, mixed with the sky by the fog transmittance
// (vx_resolve).
//
// One wave per 8x8 tile, one lane per pixel, as the march. Each lane re-derives its hit
// from the key (hit_t, block_at), gathers the 3x3 light/occlusion neighbourhood in front
// of the face (gather_face inside the chunk, world_sample across chunks), and samples the
// caller's atlas. The configuration and its f32 steps are the oracle's (oracle/
// vxresolve.inc); the elementary functions are the library's own (log2, exp2, pow =
// exp2(y log2 x), exp = wave_exp). Water and glass pixels belong to lanes 1 and 2: lane 0
// leaves them at alpha 0.
#include <cstring>

#include "ctx.hpp"

namespace vxd {

constexpr uint32_t FLAG_AO = 8u, FLAG_TEX_VARIATION = 16384u, SAND_ID = 4u, LIGHT_OPEN_SKY = 0xF000u,
                   ATTR_HAS_FOLIAGE = 4u;
constexpr float AO_STRENGTH = 0.25f, SKY_SPEC_GAIN = 0.5f, WET_BAND = 1.5f, WET_DARK = 0.62f, SKY_GRAD_EXP = 0.55f;
constexpr uint32_t PERM_NONE = 0u, PERM_D4 = 2u;

// Resolve switches (pipeline-key overrides, src/render/mod.rs:2355-2689).
constexpr uint32_t RS_SKY_SPECULAR = 1u, RS_LIGHTING_REPAIR = 2u, RS_SKY_TINT = 4u, RS_SKY_COOL = 8u, RS_SKY_LOOK = 16u,
                   RS_LIGHT_RGB = 32u, RS_EMITTER_GATHER = 64u, RS_SHORE_WET = 128u, RS_TEX_VARIATION = 256u,
                   RS_GLASS = 512u, RS_WATER_SEC = 1024u, RS_SHORE_FOAM = 2048u, RS_FLAT_SECONDARY = 4096u,
                   RS_DISTANT = 8192u, RS_TINT = 16384u, RS_TINT_BALANCE = 32768u, RS_PROBE_CUBE = 65536u,
                   RS_PROBE_BOUNCE = 131072u, RS_PROBE_SUN = 262144u, RS_PROBE_SUN_HIGH = 524288u,
                   RS_CAUSTICS = 1048576u, RS_PROBE_TAP = 4194304u, RS_PROBE_AMBIENT = 8388608u, RS_SOFT_SHADOWS = 2097152u;

struct ResolveK {
    Pools p;
    FrameK f;
    float fog_density, fog_falloff, ambient, fog_scatter, fog_g, fog_height, mip_bias, haze_warm, zenith_deep;
    uint32_t bits;  // RS_*
    const uint32_t *__restrict__ atlas;  // RGBA8 texels
    uint32_t size, layers, mips;
    const uint32_t *__restrict__ faces;
    uint32_t n_faces;
    const unsigned long long *__restrict__ vis;
    const uint16_t *__restrict__ sun;
    const uint32_t *__restrict__ heat;
    float4 *__restrict__ out;
    // lanes 1 and 2: the march switches their traces read, and the water_sec legs
    // (2x2 blocks, hw x hh texels of f16 RGBA: refract rgb + depth, reflect rgb + hit)
    Spec s;
    ushort4 *__restrict__ sec_r;
    ushort4 *__restrict__ sec_m;
    uint32_t hw, hh;
    float water_absorb;
    // distant terrain shadows: the shaft field's result slice and the frame's window
    const float *__restrict__ shaft;
    float shaft_origin[2], shaft_texel, shaft_soft;
    // the biome tint
    float tint_strength, tint_freq_t, tint_freq_h;
    int32_t tint_seed;
    // the ambient probe field (RGBA16F texels)
    const ushort4 *__restrict__ probe;
    // soft shadows: shadow_horizontal's RG16F output and shadow_vertical's R8 mask (the
    // primary sun visibility the resolve reads), or null without SPEC_SOFT_SHADOWS
    ushort2 *__restrict__ soft_h;
    uint8_t *__restrict__ soft_mask;
};

// ---- elementary functions (the oracle's vxo_log2 / vxo_exp2 / vxo_pow) ----
__device__ __forceinline__ float r_log2(float x) {
    if (x != x || x < 0.0f) return __uint_as_float(0x7FC00000u);
    if (x == 0.0f) return -__builtin_inff();
    if (__builtin_isinf(x)) return x;
    uint32_t u = __float_as_uint(x);
    int32_t e = 0;
    if (u < 0x00800000u) {
        x = x * 0x1p24f;
        u = __float_as_uint(x);
        e = -24;
    }
    e += (int32_t)(u >> 23) - 127;
    float m = __uint_as_float((u & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 0x1.3b13b2p-4f;
    p = p * s2 + 0x1.745d18p-4f;
    p = p * s2 + 0x1.c71c72p-4f;
    p = p * s2 + 0x1.24924ap-3f;
    p = p * s2 + 0x1.99999ap-3f;
    p = p * s2 + 0x1.555556p-2f;
    p = p * s2 + 1.0f;
    const float ln_m = 2.0f * s * p;
    return (float)e + ln_m * 0x1.715476p0f;
}

__device__ __forceinline__ float r_exp2(float x) {
    if (x != x) return x;
    if (x < -150.0f) return 0.0f;
    if (x > 128.0f) return __builtin_inff();
    const float k = rintf(x);
    const float r = (x - k) * 0x1.62e43p-1f;
    float e = 0x1.a01a02p-13f;
    e = e * r + 0x1.6c16c2p-10f;
    e = e * r + 0x1.111112p-7f;
    e = e * r + 0x1.555556p-5f;
    e = e * r + 0x1.555556p-3f;
    e = e * r + 0.5f;
    e = e * r + 1.0f;
    e = e * r + 1.0f;
    return ldexpf(e, (int)k);
}

__device__ __forceinline__ float r_pow(float x, float y) { return r_exp2(y * r_log2(x)); }
__device__ __forceinline__ float r_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// ---- distant terrain shadows: the shaft field (shaft.wgsl:1-24, common.wgsl:1557-1580) ----
// The buffer holds three 512 x 512 f32 slices: 0 and 1 the scan's ping-pong, 2 the heights
// (SHAFT_HEIGHTS); pass k reads the heights (k = 0) or slice (k - 1) & 1 and writes slice
// k & 1, so the eighth pass leaves the result in slice 1 (SHAFT_RESULT).
constexpr uint32_t SHAFT_DIM = 512u, SHAFT_SLICE = SHAFT_DIM * SHAFT_DIM, SHAFT_STEPS = 8u;

// shaft_sample (common.wgsl:1557-1569)
__device__ __forceinline__ float shaft_sample(const float *__restrict__ buf, float tx, float tz) {
    const float hi = (float)SHAFT_DIM - 1.0f;
    const float cx = fminf(fmaxf(tx, 0.0f), hi), cz = fminf(fmaxf(tz, 0.0f), hi);
    const float f0x = floorf(cx), f0z = floorf(cz);
    const float frx = cx - f0x, frz = cz - f0z;
    const uint32_t i0x = (uint32_t)f0x, i0z = (uint32_t)f0z;
    const uint32_t i1x = min(i0x + 1u, SHAFT_DIM - 1u), i1z = min(i0z + 1u, SHAFT_DIM - 1u);
    const float a = buf[i0z * SHAFT_DIM + i0x], b = buf[i0z * SHAFT_DIM + i1x];
    const float c0 = buf[i1z * SHAFT_DIM + i0x], d = buf[i1z * SHAFT_DIM + i1x];
    return r_mix(r_mix(a, b, frx), r_mix(c0, d, frx), frz);
}

// shaft_scan (shaft.wgsl:3-24), pass k: 16 x 16 texels a block over the 512 x 512 slice.
__global__ __launch_bounds__(256) void k_shaft_scan(float *__restrict__ shaft, uint32_t k, float sx, float sy,
                                                    float sz, float texel) {
    const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u), z = blockIdx.y * 16u + (threadIdx.x >> 4);
    if (x >= SHAFT_DIM || z >= SHAFT_DIM) return;
    const float *src = shaft + (k > 0u ? ((k - 1u) & 1u) * SHAFT_SLICE : 2u * SHAFT_SLICE);
    float *dst = shaft + (k & 1u) * SHAFT_SLICE;
    const float span = (float)(1u << k);
    const float horiz = fmaxf(sqrtf(sx * sx + sz * sz), 1e-3f);
    const float dx = sx / horiz, dz = sz / horiz;
    const float climb = span * texel * (sy / horiz);
    const float here = src[z * SHAFT_DIM + x];
    const float carried = shaft_sample(src, (float)x + dx * span, (float)z + dz * span) - climb;
    dst[z * SHAFT_DIM + x] = fmaxf(here, carried);
}
__device__ __forceinline__ float r_clamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

__device__ __forceinline__ float srgb_decode(uint32_t u) {
    const float c = (float)u / 255.0f;
    if (c <= 0.04045f) return c / 12.92f;
    return r_pow((c + 0.055f) / 1.055f, 2.4f);
}

// ---- atlas (textureSampleLevel with the atlas sampler, as the oracle restates it) ----
__device__ __forceinline__ void atlas_texel(const ResolveK &k, uint32_t level, uint32_t layer, int32_t x, int32_t y,
                                            float out[4]) {
    size_t off = 0;
    for (uint32_t l = 0; l < level; l++) off += (size_t)k.layers * (k.size >> l) * (k.size >> l);
    const uint32_t dim = k.size >> level;
    const uint32_t w = k.atlas[off + ((size_t)layer * dim + (uint32_t)y) * dim + (uint32_t)x];
    out[0] = srgb_decode(w & 0xFFu);
    out[1] = srgb_decode((w >> 8) & 0xFFu);
    out[2] = srgb_decode((w >> 16) & 0xFFu);
    out[3] = (float)(w >> 24) / 255.0f;
}

__device__ __forceinline__ int32_t clamp_texel(int32_t i, int32_t dim) { return i < 0 ? 0 : (i >= dim ? dim - 1 : i); }

__device__ inline void atlas_bilinear(const ResolveK &k, uint32_t level, uint32_t layer, const float uv[2],
                                      float out[4]) {
    const int32_t dim = (int32_t)(k.size >> level);
    const float x = uv[0] * (float)dim - 0.5f, y = uv[1] * (float)dim - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const int32_t x0 = clamp_texel((int32_t)fx, dim), x1 = clamp_texel((int32_t)fx + 1, dim);
    const int32_t y0 = clamp_texel((int32_t)fy, dim), y1 = clamp_texel((int32_t)fy + 1, dim);
    float t00[4], t10[4], t01[4], t11[4];
    atlas_texel(k, level, layer, x0, y0, t00);
    atlas_texel(k, level, layer, x1, y0, t10);
    atlas_texel(k, level, layer, x0, y1, t01);
    atlas_texel(k, level, layer, x1, y1, t11);
#pragma unroll
    for (int q = 0; q < 4; q++) out[q] = r_mix(r_mix(t00[q], t10[q], ax), r_mix(t01[q], t11[q], ax), ay);
}

__device__ inline void atlas_sample(const ResolveK &k, uint32_t layer, const float uv[2], float lod, float out[4]) {
    if (layer >= k.layers) {
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return;
    }
    if (!(lod > 0.0f)) {
        const int32_t dim = (int32_t)k.size;
        const int32_t x = clamp_texel((int32_t)floorf(uv[0] * (float)dim), dim);
        const int32_t y = clamp_texel((int32_t)floorf(uv[1] * (float)dim), dim);
        atlas_texel(k, 0, layer, x, y, out);
        return;
    }
    const float top = (float)(k.mips - 1u);
    const float l = fminf(lod, top);
    const float l0 = floorf(l);
    const float fr = l - l0;
    const uint32_t i0 = (uint32_t)l0, i1 = i0 + 1u < k.mips ? i0 + 1u : i0;
    float c0[4], c1[4];
    atlas_bilinear(k, i0, layer, uv, c0);
    atlas_bilinear(k, i1, layer, uv, c1);
#pragma unroll
    for (int q = 0; q < 4; q++) out[q] = r_mix(c0[q], c1[q], fr);
}

// ---- sky (common.wgsl:1351-1413) ----
__device__ __forceinline__ float hg_gain(float cos_t, float g) {
    const float g2 = g * g;
    const float d = 1.0f + g2 - 2.0f * g * cos_t;
    return (1.0f - g2) / r_pow(fmaxf(d, 1e-4f), 1.5f);
}

__device__ inline void sky_base(const ResolveK &k, const float rd[3], float shaft, float c[3]) {
    constexpr float SCATTER[3] = {1.0000f, 0.4770f, 0.1474f};
    constexpr float NH[3] = {0.0049f, 0.0060f, 0.0134f}, DH[3] = {0.4480f, 0.6210f, 0.8900f};
    constexpr float NZ[3] = {0.0008f, 0.0015f, 0.0039f}, DZ[3] = {0.0637f, 0.2140f, 0.8276f};
    constexpr float CH[3] = {0.4074f, 0.6057f, 0.9150f}, CZ[3] = {0.0385f, 0.1892f, 0.8600f};
    constexpr float SUN_C[3] = {1.0f, 0.8276f, 0.5225f}, DEEP[3] = {0.80f, 0.90f, 1.16f};
    const bool cool = (k.bits & RS_SKY_COOL) != 0u, look = (k.bits & RS_SKY_LOOK) != 0u;
    const float day = k.f.daylight;
    const float up = r_clamp(rd[1], 0.0f, 1.0f);
    const float gp = r_pow(up, SKY_GRAD_EXP);
    const float sd = dot3o(rd, k.f.sun);
    const float disc = wgsl_smoothstep(0.9975f, 0.9990f, sd);
    const float lobe_k = hg_gain(sd, k.fog_g) * k.fog_scatter * k.f.daylight;
    float band = 0.0f, toward = 0.0f, deep_t = 0.0f;
    const bool warm = look && k.haze_warm > 0.0f, zen = look && k.zenith_deep > 0.0f;
    if (warm) {
        band = r_pow(1.0f - up, 3.0f);
        toward = 0.35f + 0.65f * r_pow(fmaxf(sd, 0.0f), 4.0f);
    }
    if (zen) deep_t = k.zenith_deep * r_pow(up, SKY_GRAD_EXP) * day;
    const float w = wgsl_smoothstep(0.0f, -0.35f, rd[1]);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const float hz = r_mix(NH[q], cool ? CH[q] : DH[q], day);
        const float zn = r_mix(NZ[q], cool ? CZ[q] : DZ[q], day);
        float v = r_mix(hz, zn, gp);
        v = v + SUN_C[q] * disc * fmaxf(day, 0.2f);
        v = v + (SCATTER[q] * lobe_k) * shaft;
        if (warm) v = v + SCATTER[q] * (band * toward * k.haze_warm * day * 0.22f);
        if (zen) v = v * r_mix(1.0f, DEEP[q], deep_t);
        c[q] = r_mix(v, hz * 0.26f, w);
    }
}

// transmittance_from (common.wgsl:1332-1342)
__device__ __forceinline__ float transmittance_from(const ResolveK &k, float y0, float s, const float rd[3]) {
    const float lambda = k.fog_falloff;
    const float x = r_clamp(lambda * s * rd[1], -60.0f, 60.0f);
    float path = s;
    if (fabsf(x) > 1e-4f) path = s * (1.0f - wave_exp(-x)) / x;
    const float rho = k.fog_density * wave_exp(-r_clamp(lambda * (y0 - k.fog_height), -60.0f, 60.0f));
    return wave_exp(-r_clamp(rho * path, 0.0f, 60.0f));
}

// ---- light (common.wgsl:977-1097, resolve.wgsl:1-90) ----
__device__ __forceinline__ float light_curve(float level) {
    const float l = level / 15.0f;
    return l * l * (0.6f + 0.4f * l);
}
__device__ __forceinline__ float light_sky(uint32_t p) { return (float)(p >> 12); }
__device__ __forceinline__ void light_block_rgb(uint32_t p, float o[3]) {
    o[0] = (float)((p >> 8) & 0xFu);
    o[1] = (float)((p >> 4) & 0xFu);
    o[2] = (float)(p & 0xFu);
}
__device__ __forceinline__ float light_block_level(uint32_t p) {
    float c[3];
    light_block_rgb(p, c);
    return fmaxf(c[0], fmaxf(c[1], c[2]));
}
__device__ __forceinline__ void light_curve3(const float lv[3], float o[3]) {
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const float l = lv[q] / 15.0f;
        o[q] = l * l * (0.6f + 0.4f * l);
    }
}

// The light fields of a chunk record (light_base, light_flags: words 18 and 19).
__device__ __forceinline__ uint2 ld_chunk_light(const Pools &p, uint32_t ci) {
    if (ci >= p.n_chunks) ci = 0;
    const uint32_t *w = reinterpret_cast<const uint32_t *>(p.chunks + ci);
    return make_uint2(w[18], w[19]);
}

__device__ inline uint32_t light_at(const Pools &p, uint2 lt, int32_t x, int32_t y, int32_t z) {
    if (lt.y & 1u) return lt.x;
    if (x < 0 || y < 0 || z < 0 || x > 63 || y > 63 || z > 63) return LIGHT_OPEN_SKY;
    const uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
    const uint32_t cell = (ux >> 2) | ((uy >> 2) << 4) | ((uz >> 2) << 8);
    const uint32_t entry = ld_brick(p, lt.x + cell);
    if (entry & UNIFORM_BIT) return entry & 0xFFFFu;
    const uint32_t vb = cell_bit(ux & 3u, uy & 3u, uz & 3u);
    const uint32_t w = ld_brick(p, entry + (vb >> 1));
    return (w >> ((vb & 1u) * 16u)) & 0xFFFFu;
}

__device__ __forceinline__ bool is_tuft(uint32_t b) { return b == 18u || b == 24u || b == 25u; }

// voxel_occludes (common.wgsl:942-975)
__device__ inline bool voxel_occludes(const Pools &p, const ChunkK &c, int32_t x, int32_t y, int32_t z) {
    if (!voxel_occupied(p, c, x, y, z)) return false;
    if (c.attr_flags & (ATTR_HAS_WATER | ATTR_HAS_FOLIAGE)) {
        const uint32_t b = block_at(p, c, (uint32_t)x, (uint32_t)y, (uint32_t)z);
        return b != WATER_ID && !is_tuft(b);
    }
    return true;
}

// world_sample (common.wgsl:1014-1032): (light word, occludes)
__device__ inline uint2 world_sample(const Pools &p, const FrameK &f, const float pt[3]) {
    bool ok;
    const uint32_t idx = grid_index(f, (int32_t)floorf(pt[0] / 64.0f), (int32_t)floorf(pt[1] / 64.0f),
                                    (int32_t)floorf(pt[2] / 64.0f), ok);
    if (!ok) return make_uint2(LIGHT_OPEN_SKY, 0u);
    const uint32_t ci = ld_grid(p, f, idx);
    if (ci == NO_CHUNK || ci >= p.n_chunks) return make_uint2(LIGHT_OPEN_SKY, 0u);
    const ChunkK c = ld_chunk(p, ci);
    const int32_t x = (int32_t)floorf((pt[0] - c.origin[0]) / c.vs), y = (int32_t)floorf((pt[1] - c.origin[1]) / c.vs),
                  z = (int32_t)floorf((pt[2] - c.origin[2]) / c.vs);
    return make_uint2(light_at(p, ld_chunk_light(p, ci), x, y, z), voxel_occludes(p, c, x, y, z) ? 1u : 0u);
}

// sky_ambient_tint and surface_sky_fill (resolve.wgsl:1-5, :56-89)
__device__ inline void surface_sky_fill(const ResolveK &k, const float n[3], float o[3]) {
    constexpr float SKY_TINT[3] = {0.6038f, 0.7084f, 1.0000f};
    constexpr float REF[3] = {0.14658824f, 0.30178431f, 0.84105882f};
    constexpr float SCATTER[3] = {1.0000f, 0.4770f, 0.1474f};
    constexpr float NH[3] = {0.0049f, 0.0060f, 0.0134f}, DH[3] = {0.4480f, 0.6210f, 0.8900f};
    constexpr float NZ[3] = {0.0008f, 0.0015f, 0.0039f}, DZ[3] = {0.0637f, 0.2140f, 0.8276f};
    if (!(k.bits & RS_SKY_TINT)) {
#pragma unroll
        for (int q = 0; q < 3; q++) o[q] = SKY_TINT[q];
    } else {
        const float day = k.f.daylight;
        const float w = n[1] < -0.5f ? 0.0f : (n[1] > 0.5f ? 0.78431373f : 0.58578900f);
        const float h = k.fog_scatter * day * 4.0f * fmaxf(dot3o(n, k.f.sun), 0.0f);
        float t[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const float c = r_mix(r_mix(NH[q], DH[q], day), r_mix(NZ[q], DZ[q], day), w) + SCATTER[q] * h;
            t[q] = SKY_TINT[q] * r_pow(c / REF[q], 2.5f);
        }
        const float luma = 0.2126f * t[0] + 0.7152f * t[1] + 0.0722f * t[2];
        const float s = 0.70721556f / fmaxf(luma, 1e-6f);
#pragma unroll
        for (int q = 0; q < 3; q++) o[q] = t[q] * s;
    }
    if (!(k.bits & RS_LIGHTING_REPAIR)) return;
    const float m = 0.60f * r_mix(0.35f, 1.0f, r_clamp(n[1] * 0.5f + 0.5f, 0.0f, 1.0f));
#pragma unroll
    for (int q = 0; q < 3; q++) o[q] = o[q] * m;
}

__device__ __forceinline__ float schlick_ground(float cos_i) {
    const float m = r_clamp(1.0f - cos_i, 0.0f, 1.0f);
    const float m2 = m * m;
    const float top = fmaxf(1.0f - 0.9f, 0.04f);
    return 0.04f + (top - 0.04f) * m2 * m2 * m;
}

__device__ __forceinline__ uint32_t perm_key(const float b[3], uint32_t nid) {
    const int32_t c0 = (int32_t)floorf(b[0]), c1 = (int32_t)floorf(b[1]), c2 = (int32_t)floorf(b[2]);
    return hash_u32(((uint32_t)c0 * 0x9e3779b9u) ^ ((uint32_t)c1 * 0x85ebca6bu) ^ ((uint32_t)c2 * 0xc2b2ae35u) ^
                    (nid * 0x27d4eb2fu));
}

// ---- biome tint (resolve.wgsl:166-287) ----
constexpr uint32_t FLAG_TINT = 128u, BIOME_PRIME_X = 501125321u, BIOME_PRIME_Y = 1136930381u;
constexpr int32_t BIOME_TEMP_SEED = 7, BIOME_HUMID_SEED = 8;
constexpr float BIOME_BAND = 0.22f, BIOME_BLEND = 0.18f;
constexpr float GRAD24_FIRST = 1.4398966328953218f, GRAD24_STEP = 0.26179938779914946f;
constexpr float GRAD8_FIRST = 1.1780972450961724f, GRAD8_STEP = 0.7853981633974483f;

__device__ __forceinline__ float r_sin(float x) {  // the oracle's vxo_wave_sin
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
    const int32_t q = (int32_t)kf & 3;
    return q == 0 ? sn : (q == 1 ? cs : (q == 2 ? -sn : -cs));
}

// simplex_grad (resolve.wgsl:180-192)
__device__ __forceinline__ float simplex_grad(uint32_t seed, uint32_t xp, uint32_t yp, float xd, float yd) {
    const int32_t hs = (int32_t)((seed ^ xp ^ yp) * 0x27d4eb2du);
    const uint32_t idx = (uint32_t)((hs ^ (hs >> 15)) & 254);
    const float angle = idx < 240u ? GRAD24_FIRST - GRAD24_STEP * (float)((idx % 48u) >> 1)
                                   : GRAD8_FIRST - GRAD8_STEP * (float)((idx - 240u) >> 1);
    return xd * wave_cos(angle) + yd * r_sin(angle);
}

// biome_fast_floor (resolve.wgsl:194-199)
__device__ __forceinline__ int32_t biome_fast_floor(float f) { return f >= 0.0f ? (int32_t)f : (int32_t)f - 1; }

// biome_noise (resolve.wgsl:201-257): 2D OpenSimplex2, three corner kernels.
__device__ inline float biome_noise(int32_t seed_i, float freq, float px, float pz) {
    const uint32_t seed = (uint32_t)seed_i;
    const float sqrt3 = 1.7320508075688772f;
    const float g2 = (3.0f - sqrt3) / 6.0f;
    const float f2 = 0.5f * (sqrt3 - 1.0f);
    float x = px * freq, y = pz * freq;
    const float s = (x + y) * f2;
    x = x + s;
    y = y + s;
    const int32_t i0 = biome_fast_floor(x), j0 = biome_fast_floor(y);
    const float xi = x - (float)i0, yi = y - (float)j0;
    const float t = (xi + yi) * g2;
    const float x0 = xi - t, y0 = yi - t;
    const uint32_t i = (uint32_t)i0 * BIOME_PRIME_X, j = (uint32_t)j0 * BIOME_PRIME_Y;
    const float a = 0.5f - x0 * x0 - y0 * y0;
    float n0 = 0.0f;
    if (a > 0.0f) n0 = (a * a) * (a * a) * simplex_grad(seed, i, j, x0, y0);
    const float c = (2.0f * (1.0f - 2.0f * g2) * (1.0f / g2 - 2.0f)) * t +
                    ((-2.0f * (1.0f - 2.0f * g2) * (1.0f - 2.0f * g2)) + a);
    float n2 = 0.0f;
    if (c > 0.0f) {
        const float x2 = x0 + (2.0f * g2 - 1.0f), y2 = y0 + (2.0f * g2 - 1.0f);
        n2 = (c * c) * (c * c) * simplex_grad(seed, i + BIOME_PRIME_X, j + BIOME_PRIME_Y, x2, y2);
    }
    const bool upper = y0 > x0;
    const float x1 = upper ? x0 + g2 : x0 + (g2 - 1.0f), y1 = upper ? y0 + (g2 - 1.0f) : y0 + g2;
    const float b = 0.5f - x1 * x1 - y1 * y1;
    float n1 = 0.0f;
    if (b > 0.0f)
        n1 = (b * b) * (b * b) * simplex_grad(seed, upper ? i : i + BIOME_PRIME_X, upper ? j + BIOME_PRIME_Y : j, x1, y1);
    return (n0 + n1 + n2) * 99.83685446303647f;
}

// biome_band_weights and biome_tint (resolve.wgsl:259-287)
__device__ inline void biome_tint(const ResolveK &k, float x, float z, float out[3]) {
    constexpr float PLAINS[3] = {1.000f, 1.000f, 1.000f}, PLAINS_BALANCED[3] = {0.935f, 1.067f, 0.892f};
    constexpr float FOREST[3] = {0.664f, 1.045f, 1.000f}, SAVANNA[3] = {1.842f, 0.935f, 0.793f};
    constexpr float DESERT[3] = {2.096f, 0.957f, 0.699f}, TAIGA[3] = {0.832f, 0.935f, 2.096f};
    constexpr float TUNDRA[3] = {0.755f, 0.893f, 2.812f};
    const float temp = biome_noise(k.tint_seed + BIOME_TEMP_SEED, k.tint_freq_t, x, z);
    const float humid = biome_noise(k.tint_seed + BIOME_HUMID_SEED, k.tint_freq_h, x, z);
    float wt[3], wh[3];
    const float n2[2] = {temp, humid};
    float *w2[2] = {wt, wh};
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const float hot = wgsl_smoothstep(BIOME_BAND - BIOME_BLEND, BIOME_BAND + BIOME_BLEND, n2[u]);
        const float cold = 1.0f - wgsl_smoothstep(-BIOME_BAND - BIOME_BLEND, -BIOME_BAND + BIOME_BLEND, n2[u]);
        w2[u][0] = cold;
        w2[u][1] = 1.0f - cold - hot;
        w2[u][2] = hot;
    }
    const bool balance = (k.bits & RS_TINT_BALANCE) != 0u;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const float pl = balance ? PLAINS_BALANCED[q] : PLAINS[q];
        const float dry = wt[0] * TUNDRA[q] + wt[1] * pl + wt[2] * DESERT[q];
        const float mid = wt[0] * TAIGA[q] + wt[1] * pl + wt[2] * SAVANNA[q];
        const float wet = wt[0] * TAIGA[q] + wt[1] * FOREST[q] + wt[2] * FOREST[q];
        out[q] = wh[0] * dry + wh[1] * mid + wh[2] * wet;
    }
}

// ---- the ambient probe field (resolve.wgsl:95-118, src/render/mod.rs:2240-2315) ----
constexpr uint32_t PROBE_XZ = 128u, PROBE_Y = 128u, PROBE_SLABS = 6u, PROBE_H = PROBE_Y * PROBE_SLABS;
constexpr size_t PROBE_TEXELS = (size_t)PROBE_XZ * PROBE_H * PROBE_XZ;

__device__ __forceinline__ void probe_texel(const ushort4 *__restrict__ t, uint32_t x, uint32_t y, uint32_t z,
                                            float out[4]) {
    const ushort4 v = t[((size_t)z * PROBE_H + y) * PROBE_XZ + x];
    out[0] = (float)__builtin_bit_cast(_Float16, v.x);
    out[1] = (float)__builtin_bit_cast(_Float16, v.y);
    out[2] = (float)__builtin_bit_cast(_Float16, v.z);
    out[3] = (float)__builtin_bit_cast(_Float16, v.w);
}

// probe_field: the slab of the face's normal at p / PROBE_SPACING, y kept half a texel
// inside the slab, sampled as the probe sampler does (linear; repeat in x and z, clamp to
// edge in y): texel coordinates uvw * size - 0.5, eight texels, weights in f32.
__device__ inline void probe_field(const ResolveK &k, const float p[3], uint32_t nid, float out[4]) {
    const float tx = p[0] * (1.0f / 4.0f), ty = p[1] * (1.0f / 4.0f), tz = p[2] * (1.0f / 4.0f);
    const float y = fminf(fmaxf(ty, 0.5f), (float)PROBE_Y - 0.5f);
    const float u = tx * (1.0f / (float)PROBE_XZ);
    const uint32_t slab = nid >= NORMAL_CROSS_A ? 3u : nid;
    const float v = ((float)slab * (float)PROBE_Y + y) * (1.0f / ((float)PROBE_Y * (float)PROBE_SLABS));
    const float w = tz * (1.0f / (float)PROBE_XZ);
    const float X = u * (float)PROBE_XZ - 0.5f, Y = v * (float)PROBE_H - 0.5f, Z = w * (float)PROBE_XZ - 0.5f;
    const float fx0 = floorf(X), fy0 = floorf(Y), fz0 = floorf(Z);
    const float ax = X - fx0, ay = Y - fy0, az = Z - fz0;
    const int32_t ix = (int32_t)fx0, iy = (int32_t)fy0, iz = (int32_t)fz0;
    const uint32_t x0 = (uint32_t)(((ix % 128) + 128) % 128), x1 = (x0 + 1u) & 127u;
    const uint32_t z0 = (uint32_t)(((iz % 128) + 128) % 128), z1 = (z0 + 1u) & 127u;
    const uint32_t y0 = (uint32_t)min(max(iy, 0), (int32_t)PROBE_H - 1);
    const uint32_t y1 = (uint32_t)min(max(iy + 1, 0), (int32_t)PROBE_H - 1);
    float c[8][4];
    probe_texel(k.probe, x0, y0, z0, c[0]);
    probe_texel(k.probe, x1, y0, z0, c[1]);
    probe_texel(k.probe, x0, y1, z0, c[2]);
    probe_texel(k.probe, x1, y1, z0, c[3]);
    probe_texel(k.probe, x0, y0, z1, c[4]);
    probe_texel(k.probe, x1, y0, z1, c[5]);
    probe_texel(k.probe, x0, y1, z1, c[6]);
    probe_texel(k.probe, x1, y1, z1, c[7]);
#pragma unroll
    for (int q = 0; q < 4; q++)
        out[q] = r_mix(r_mix(r_mix(c[0][q], c[1][q], ax), r_mix(c[2][q], c[3][q], ax), ay),
                       r_mix(r_mix(c[4][q], c[5][q], ax), r_mix(c[6][q], c[7][q], ax), ay), az);
}

// probe_tap (resolve.wgsl:120-127, one tap): the field sampled at fract(p / 512) across the
// whole texture, all six slabs included, with the linear sampler (clamp to edge on every axis).
__device__ inline void probe_tap(const ResolveK &k, const float p[3], float out[3]) {
    float uvw[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const float s = p[q] * (1.0f / 512.0f) + 0.0f;
        uvw[q] = s - floorf(s);
    }
    const float X = uvw[0] * (float)PROBE_XZ - 0.5f, Y = uvw[1] * (float)PROBE_H - 0.5f,
                Z = uvw[2] * (float)PROBE_XZ - 0.5f;
    const float fx0 = floorf(X), fy0 = floorf(Y), fz0 = floorf(Z);
    const float ax = X - fx0, ay = Y - fy0, az = Z - fz0;
    const int32_t ix = (int32_t)fx0, iy = (int32_t)fy0, iz = (int32_t)fz0;
    const int32_t xz_max = (int32_t)PROBE_XZ - 1, h_max = (int32_t)PROBE_H - 1;
    const uint32_t x0 = (uint32_t)min(max(ix, 0), xz_max), x1 = (uint32_t)min(max(ix + 1, 0), xz_max);
    const uint32_t y0 = (uint32_t)min(max(iy, 0), h_max), y1 = (uint32_t)min(max(iy + 1, 0), h_max);
    const uint32_t z0 = (uint32_t)min(max(iz, 0), xz_max), z1 = (uint32_t)min(max(iz + 1, 0), xz_max);
    float c[8][4];
    probe_texel(k.probe, x0, y0, z0, c[0]);
    probe_texel(k.probe, x1, y0, z0, c[1]);
    probe_texel(k.probe, x0, y1, z0, c[2]);
    probe_texel(k.probe, x1, y1, z0, c[3]);
    probe_texel(k.probe, x0, y0, z1, c[4]);
    probe_texel(k.probe, x1, y0, z1, c[5]);
    probe_texel(k.probe, x0, y1, z1, c[6]);
    probe_texel(k.probe, x1, y1, z1, c[7]);
#pragma unroll
    for (int q = 0; q < 3; q++)
        out[q] = 1.0f * r_mix(r_mix(r_mix(c[0][q], c[1][q], ax), r_mix(c[2][q], c[3][q], ax), ay),
                              r_mix(r_mix(c[4][q], c[5][q], ax), r_mix(c[6][q], c[7][q], ax), ay), az);
}

// terrain_shade_beyond (common.wgsl:1571-1580): the scanned horizon at p + sun_dir * d
// (texel centres at origin + (i + 0.5) texel), ramped over +-shaft_soft around its y.
__device__ inline float terrain_shade_beyond(const ResolveK &k, const float p[3], float d) {
    float q[3];
#pragma unroll
    for (int i = 0; i < 3; i++) q[i] = p[i] + k.f.sun[i] * d;
    const float tx = (q[0] - k.shaft_origin[0]) / k.shaft_texel - 0.5f;
    const float tz = (q[2] - k.shaft_origin[1]) / k.shaft_texel - 0.5f;
    const float e = shaft_sample(k.shaft, tx, tz);
    return wgsl_smoothstep(e - k.shaft_soft, e + k.shaft_soft, q[1]);
}


// shade_hit_compact_cached / shade_hit_legacy_cached (resolve.wgsl:1537-2059) with smooth
// light and do_spec, for lane 0; shade_hit_legacy (:289-586) for the secondary legs, whose
// sun visibility lit_word the caller takes from shadow_ray, without do_spec, and with flat
// light (the light word in front of the face, no occlusion) when smooth is false. Faces
// only (not foliage quads).
__device__ inline void shade_opaque(const ResolveK &k, uint32_t ci, uint32_t vx, uint32_t vy, uint32_t vz, uint32_t id,
                                    uint32_t nid, const float hit[3], const float rd[3], float dist, float lit_word,
                                    float shadow_dist, bool smooth, bool do_spec, float out[3]) {
    const Pools &p = k.p;
    const FrameK &f = k.f;
    const ChunkK ch = ld_chunk(p, ci);
    const uint2 clight = ld_chunk_light(p, ci);
    const float vs = ch.vs;
    float n[3];
    normal_of(nid, n);
    float spec_pre[3] = {0.0f, 0.0f, 0.0f};
    if ((k.bits & RS_SKY_SPECULAR) && do_spec) {
        float m[3], sb[3];
        wgsl_reflect(rd, n, m);
        sky_base(k, m, 1.0f, sb);
        const float nr[3] = {-rd[0], -rd[1], -rd[2]};
        const float g = schlick_ground(r_clamp(dot3o(n, nr), 0.0f, 1.0f));
#pragma unroll
        for (int q = 0; q < 3; q++) spec_pre[q] = sb[q] * g;
    }
    const float grazing = fmaxf(fabsf(dot3o(n, rd)), 0.25f);
    const float footprint = fmaxf(dist, 1e-3f) * 2.0f * f.tan_half_fov / (float)f.res[1] * 16.0f / vs / grazing;
    const float lod = r_clamp(r_log2(fmaxf(footprint, 1e-4f)) + k.mip_bias, 0.0f, 3.0f);
    const uint32_t axis = nid >> 1;
    const uint32_t v[3] = {vx, vy, vz};
    float box_min[3], local[3];
#pragma unroll
    for (int q = 0; q < 3; q++) box_min[q] = ch.origin[q] + (float)v[q] * vs;
#pragma unroll
    for (int q = 0; q < 3; q++) local[q] = r_clamp((hit[q] - box_min[q]) / vs, 0.0f, 1.0f);
    float fa, fb;
    int32_t ta[3] = {0, 0, 0}, tb[3] = {0, 0, 0};
    if (axis == 0u) { fa = local[1]; fb = local[2]; ta[1] = 1; tb[2] = 1; }
    else if (axis == 1u) { fa = local[0]; fb = local[2]; ta[0] = 1; tb[2] = 1; }
    else { fa = local[0]; fb = local[1]; ta[0] = 1; tb[1] = 1; }
    float sky_tint[3];
    surface_sky_fill(k, n, sky_tint);
    float ao = 1.0f, sky_sm, amb3[3];
    if (!smooth) {
        // flat light: the light word of the air voxel in front of the face
        const int32_t a0 = (int32_t)v[0] + (int32_t)n[0], a1 = (int32_t)v[1] + (int32_t)n[1],
                      a2 = (int32_t)v[2] + (int32_t)n[2];
        uint32_t packed1;
        if (a0 >= 0 && a1 >= 0 && a2 >= 0 && a0 <= 63 && a1 <= 63 && a2 <= 63) {
            packed1 = light_at(p, clight, a0, a1, a2);
        } else {
            float pt[3];
#pragma unroll
            for (int q = 0; q < 3; q++) pt[q] = (box_min[q] + 0.5f * vs) + n[q] * vs;
            packed1 = world_sample(p, f, pt).x;
        }
        const float sk1 = light_curve(light_sky(packed1)) * f.daylight;
        sky_sm = sk1;
        if (k.bits & RS_LIGHT_RGB) {
            float lv[3], bl[3];
            light_block_rgb(packed1, lv);
            light_curve3(lv, bl);
            const float bk = fmaxf(bl[0], fmaxf(bl[1], bl[2]));
            const float t = bk / fmaxf(sk1 + bk, 1e-4f);
            const float m = fmaxf(sk1, bk);
#pragma unroll
            for (int q = 0; q < 3; q++) amb3[q] = r_mix(sky_tint[q], bl[q] / fmaxf(bk, 1e-4f), t) * m;
        } else {
            constexpr float BLOCK_TINT1[3] = {1.0000f, 0.6500f, 0.3200f};
            const float bk1 = light_curve(light_block_level(packed1));
#pragma unroll
            for (int q = 0; q < 3; q++)
                amb3[q] = r_mix(sky_tint[q], BLOCK_TINT1[q], bk1 / fmaxf(sk1 + bk1, 1e-4f)) * fmaxf(sk1, bk1);
        }
    } else {
        int32_t air_v[3];
        bool local_only = true;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            air_v[q] = (int32_t)v[q] + (int32_t)n[q];
            const int32_t span = abs(ta[q]) + abs(tb[q]);
            if (!(air_v[q] - span >= 0 && air_v[q] + span <= 63)) local_only = false;
        }
        float sky_n[9], blk_n[9], open_n[9];
        uint32_t light_w[9];
        uint32_t any_blk = 0u;
        const bool has_water = (ch.attr_flags & ATTR_HAS_WATER) != 0u;
        for (uint32_t i = 0; i < 9u; i++) {
            const int32_t da = (int32_t)(i % 3u) - 1, db = (int32_t)(i / 3u) - 1;
            uint2 nb;
            if (local_only) {  // gather_face (common.wgsl:1034-1097)
                const int32_t x = air_v[0] + ta[0] * da + tb[0] * db, y = air_v[1] + ta[1] * da + tb[1] * db,
                              z = air_v[2] + ta[2] * da + tb[2] * db;
                bool solid = voxel_occupied(p, ch, x, y, z);
                if (solid && has_water) solid = block_at(p, ch, (uint32_t)x, (uint32_t)y, (uint32_t)z) != WATER_ID;
                nb = make_uint2((clight.y & 1u) ? clight.x : light_at(p, clight, x, y, z), solid ? 1u : 0u);
            } else {
                float pt[3];
#pragma unroll
                for (int q = 0; q < 3; q++)
                    pt[q] = ((box_min[q] + 0.5f * vs) + n[q] * vs) + ((float)ta[q] * vs) * ((float)(i % 3u) - 1.0f) +
                            ((float)tb[q] * vs) * ((float)(i / 3u) - 1.0f);
                nb = world_sample(p, f, pt);
            }
            light_w[i] = nb.x;
            sky_n[i] = light_curve(light_sky(nb.x)) * f.daylight;
            blk_n[i] = 0.0f;
            if (k.bits & RS_LIGHT_RGB) {
                if (k.bits & RS_EMITTER_GATHER) any_blk = any_blk | (nb.x & 0xFFFu);
            } else {
                blk_n[i] = light_curve(light_block_level(nb.x));
            }
            open_n[i] = 1.0f - (float)nb.y;
        }
        const bool rgb = (k.bits & RS_LIGHT_RGB) != 0u, gather = rgb && (k.bits & RS_EMITTER_GATHER) && any_blk != 0u;
        float blk4[3] = {0.0f, 0.0f, 0.0f};
        if (gather) {
            float lv[3];
            light_block_rgb(light_w[4], lv);
            light_curve3(lv, blk4);
        }
        const bool ao_on = (f.flags & FLAG_AO) != 0u;
        float amb_c[4][3], sky_c[4], ao_c[4];
#pragma unroll
        for (uint32_t kk = 0; kk < 4u; kk++) {
            const uint32_t ia = (uint32_t)(4 + ((int32_t)(kk & 1u) * 2 - 1));
            const uint32_t ib = (uint32_t)(4 + ((int32_t)(kk >> 1) * 2 - 1) * 3);
            const uint32_t ic = ia + ib - 4u;
            const float inv = 1.0f / (1.0f + open_n[ia] + open_n[ib] + open_n[ic]);
            const float sk = (sky_n[4] + sky_n[ia] * open_n[ia] + sky_n[ib] * open_n[ib] + sky_n[ic] * open_n[ic]) * inv;
            sky_c[kk] = sk;
            if (rgb) {
                if (gather) {
                    float ba[3], bb[3], bc[3], lv[3], bkv[3];
                    light_block_rgb(light_w[ia], lv);
                    light_curve3(lv, ba);
                    light_block_rgb(light_w[ib], lv);
                    light_curve3(lv, bb);
                    light_block_rgb(light_w[ic], lv);
                    light_curve3(lv, bc);
#pragma unroll
                    for (int q = 0; q < 3; q++)
                        bkv[q] = (blk4[q] + ba[q] * open_n[ia] + bb[q] * open_n[ib] + bc[q] * open_n[ic]) * inv;
                    const float bk = fmaxf(bkv[0], fmaxf(bkv[1], bkv[2]));
                    const float t = bk / fmaxf(sk + bk, 1e-4f);
                    const float m = fmaxf(sk, bk);
#pragma unroll
                    for (int q = 0; q < 3; q++) amb_c[kk][q] = r_mix(sky_tint[q], bkv[q] / fmaxf(bk, 1e-4f), t) * m;
                } else {
#pragma unroll
                    for (int q = 0; q < 3; q++) amb_c[kk][q] = sky_tint[q] * sk;
                }
            } else {
                constexpr float BLOCK_TINT[3] = {1.0000f, 0.6500f, 0.3200f};
                const float bk = (blk_n[4] + blk_n[ia] * open_n[ia] + blk_n[ib] * open_n[ib] + blk_n[ic] * open_n[ic]) * inv;
                const float t = bk / fmaxf(sk + bk, 1e-4f);
                const float m = fmaxf(sk, bk);
#pragma unroll
                for (int q = 0; q < 3; q++) amb_c[kk][q] = r_mix(sky_tint[q], BLOCK_TINT[q], t) * m;
            }
            const float s1 = 1.0f - open_n[ia];
            const float s2 = 1.0f - open_n[ib];
            const float occ = (s1 > 0.5f && s2 > 0.5f) ? 3.0f : s1 + s2 + 1.0f - open_n[ic];
            ao_c[kk] = ao_on ? 1.0f - occ * AO_STRENGTH : 1.0f;
        }
        ao = r_mix(r_mix(ao_c[0], ao_c[1], fa), r_mix(ao_c[2], ao_c[3], fa), fb);
        sky_sm = r_mix(r_mix(sky_c[0], sky_c[1], fa), r_mix(sky_c[2], sky_c[3], fa), fb);
#pragma unroll
        for (int q = 0; q < 3; q++)
            amb3[q] = r_mix(r_mix(amb_c[0][q], amb_c[1][q], fa), r_mix(amb_c[2][q], amb_c[3][q], fa), fb);
    }
    // SPEC_PROBE_AMBIENT: the tap replaces the gathered ambient; SPEC_PROBE_TAP scales it
    const bool probe_amb = (k.bits & RS_PROBE_AMBIENT) != 0u;
    if (k.bits & (RS_PROBE_AMBIENT | RS_PROBE_TAP)) {
        float tap[3];
        probe_tap(k, hit, tap);
#pragma unroll
        for (int q = 0; q < 3; q++) amb3[q] = probe_amb ? tap[q] : amb3[q] * tap[q];
    }
    float direct = 0.0f;
    const float ndl = dot3o(n, f.sun);
    if (ndl > 0.0f && f.daylight > 0.01f) {
        float lit = 1.0f;
        if (f.flags & VX_FLAG_SHADOWS) {
            lit = lit_word;
            // terrain beyond the march's reach (resolve.wgsl:516-518, :1729-1731)
            if ((k.bits & RS_DISTANT) && (f.flags & VX_FLAG_DISTANT_SHADOWS) && lit > 0.0f)
                lit = lit * terrain_shade_beyond(k, hit, shadow_dist);
        }
        direct = lit * ndl * sky_sm;
        // caustics under the sea surface (resolve.wgsl:530-549): the two longest waves'
        // slopes at the surface point above along the sun, squared into light sheets
        const float cdepth = f.sea_level - hit[1];
        if ((k.bits & RS_CAUSTICS) && cdepth > 0.0f && cdepth < 8.0f && direct > 0.0f) {
            const float run = cdepth / fmaxf(f.sun[1], 1e-3f);
            const float sp[2] = {hit[0] - f.sun[0] * run, hit[2] - f.sun[2] * run};
            const float a = f.wave_amp, l0 = f.wave_scale, ph = f.time * f.wave_speed * WAVE_TAU;
            const float zero[2] = {0.0f, 0.0f};
            float g0[3], g1[3];
            wave_octave(sp, 0.913089f, 0.407760f, l0, a * 0.42f, ph, 0.35f, zero, 0.0f, false, g0);
            wave_octave(sp, -0.388685f, 0.921371f, l0 / 2.17f, a * 0.27f, ph * 1.4731f, 0.35f, zero, 0.0f, false,
                        g1);
            const float r0 = g0[0] * g0[0] + g0[1] * g0[1], r1 = g1[0] * g1[0] + g1[1] * g1[1];
            const float sheets = r0 * r1 * 4000.0f;
            direct = direct * (1.0f + sheets * wave_exp(-cdepth * 0.30f));
        }
    }
    constexpr float FACE_SHADE[8] = {0.62f, 0.62f, 0.50f, 1.00f, 0.80f, 0.80f, 0.90f, 0.90f};
    const bool repair = (k.bits & RS_LIGHTING_REPAIR) != 0u;
    // the probe field at hit + n * PROBE_NORMAL_BIAS (resolve.wgsl:552-576)
    float probe[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    if (!probe_amb && (k.bits & (RS_PROBE_CUBE | RS_PROBE_BOUNCE))) {
        float pp[3];
#pragma unroll
        for (int q = 0; q < 3; q++) pp[q] = hit[q] + n[q] * 2.0f;
        probe_field(k, pp, nid, probe);
    }
    float shade = probe_amb || repair ? 1.0f : FACE_SHADE[nid];
    if (!probe_amb && (k.bits & RS_PROBE_CUBE)) shade = shade * probe[3];
    const float sun_gain = (k.bits & RS_PROBE_SUN_HIGH) ? 0.25f : 0.10f;
    const float sun_k = direct * (repair ? 0.78f : 0.52f);
    constexpr float SUN_RGB[3] = {1.0000f, 0.8469f, 0.5647f}, FLOOR_RGB[3] = {0.85f, 0.90f, 1.00f};
    float light[3], spec_rgb[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const float ambient_rgb = amb3[q] * 0.50f;
        const float sun_rgb = SUN_RGB[q] * sun_k;
        float floor_rgb = 0.0f;
        if (!probe_amb) {
            floor_rgb = FLOOR_RGB[q] * k.ambient;
            if (k.bits & RS_PROBE_BOUNCE) floor_rgb = floor_rgb * probe[q];
            if (k.bits & RS_PROBE_SUN) floor_rgb = floor_rgb + FLOOR_RGB[q] * probe[q] * (sun_gain * sky_sm);
        }
        spec_rgb[q] = spec_pre[q] * (sky_sm * SKY_SPEC_GAIN);
        light[q] = repair ? ambient_rgb * ao * shade + sun_rgb + floor_rgb : ((ambient_rgb + sun_rgb) * ao + floor_rgb) * shade;
    }
    const uint32_t fi = id * 8u + nid;
    const uint32_t face_word = fi < k.n_faces ? k.faces[fi] : 0u;
    float uv[2];
    if (axis == 0u) { uv[0] = local[2]; uv[1] = 1.0f - local[1]; }
    else if (axis == 1u) { uv[0] = local[0]; uv[1] = local[2]; }
    else { uv[0] = local[0]; uv[1] = 1.0f - local[1]; }
    if (nid == 1u || nid == 4u) uv[0] = 1.0f - uv[0];
    if ((k.bits & RS_TEX_VARIATION) && (f.flags & FLAG_TEX_VARIATION)) {  // permute_uv (resolve.wgsl:157-164)
        const uint32_t pclass = face_word >> 16, h = perm_key(box_min, nid);
        const bool d4 = pclass == PERM_D4;
        float p0 = uv[0], p1 = uv[1];
        if (d4 && (h & 1u)) { p0 = uv[1]; p1 = uv[0]; }
        if (pclass != PERM_NONE && (h & 2u)) p0 = 1.0f - p0;
        if (d4 && (h & 4u)) p1 = 1.0f - p1;
        uv[0] = p0;
        uv[1] = p1;
    }
    float texel[4];
    atlas_sample(k, face_word & 0xFFFFu, uv, lod, texel);
    float albedo[3] = {texel[0], texel[1], texel[2]};
    // the biome tint, weighted by the texel's alpha (resolve.wgsl:337-343; faces only)
    if ((k.bits & RS_TINT) && (f.flags & FLAG_TINT) && texel[3] > 0.0f) {
        const float kt = texel[3] * k.tint_strength;
        float tint[3];
        biome_tint(k, hit[0], hit[2], tint);
#pragma unroll
        for (int q = 0; q < 3; q++) albedo[q] = albedo[q] * r_mix(1.0f, tint[q], kt);
    }
    if ((k.bits & RS_SHORE_WET) && id == SAND_ID) {
        const float wet = 1.0f - wgsl_smoothstep(f.sea_level, f.sea_level + WET_BAND, hit[1]);
        const float m = r_mix(1.0f, WET_DARK, wet);
#pragma unroll
        for (int q = 0; q < 3; q++) albedo[q] = albedo[q] * m;
    }
    if (repair && id == GLOWSTONE_ID) {
        constexpr float GLOW[3] = {1.0f, 0.7111f, 0.3024f};
#pragma unroll
        for (int q = 0; q < 3; q++) out[q] = albedo[q] * (GLOW[q] * 1.5f + light[q] * 0.15f);
        return;
    }
#pragma unroll
    for (int q = 0; q < 3; q++) out[q] = albedo[q] * (light[q] * 1.0f) + spec_rgb[q];
}

// ---- Lanes 1 and 2: glass and water pixels (resolve.wgsl:1029-1310, :1379-1466) ---------
// As the oracle restates them: shade_glass traces the transmit leg (and the reflect leg
// with SPEC_GLASS_REFLECT) and shades its hit (shade_water on water); shade_water builds
// water_ctx, traces the refract and reflect legs and composes them; with SPEC_WATER_SEC the
// legs of each 2x2 block come from k_water_sec (stored as f16) where water_share_safe
// holds. Leg hits: shade_hit_legacy with the sun visibility of shadow_ray, no sky specular,
// flat light under SPEC_FLAT_SECONDARY.
constexpr float WATER_F0 = 0.02f, GLASS_F0 = 0.0426f, GLINT_POWER = 320.0f, GLINT_GAIN = 2.5f, FOAM_DEPTH = 2.5f,
                FOAM_LO = 0.45f, FOAM_HI = 0.75f, FOAM_MAX = 0.85f, FOAM_RATE = 0.55f;
constexpr uint32_t WATER_SEC_SPAN = 2u;


// face_uv (resolve.wgsl:129-147)
__device__ __forceinline__ void face_uv_of(uint32_t nid, const float local[3], float uv[2]) {
    const uint32_t axis = nid >> 1;
    if (axis == 0u) { uv[0] = local[2]; uv[1] = 1.0f - local[1]; }
    else if (axis == 1u) { uv[0] = local[0]; uv[1] = local[2]; }
    else { uv[0] = local[0]; uv[1] = 1.0f - local[1]; }
    if (nid == 1u || nid == 4u) uv[0] = 1.0f - uv[0];
}

__device__ __forceinline__ float schlick_water(float cos_i) {
    const float m = r_clamp(1.0f - cos_i, 0.0f, 1.0f);
    const float m2 = m * m;
    return WATER_F0 + (1.0f - WATER_F0) * m2 * m2 * m;
}
__device__ __forceinline__ float schlick_glass(float cos_i) {
    const float m = r_clamp(1.0f - cos_i, 0.0f, 1.0f);
    const float m2 = m * m;
    return GLASS_F0 + (1.0f - GLASS_F0) * m2 * m2 * m;
}

__device__ __forceinline__ float medium_light(const ResolveK &k, uint32_t packed) {
    const float sky_l = light_curve(light_sky(packed)) * k.f.daylight;
    const float blk_l = light_curve(light_block_level(packed));
    return fmaxf(sky_l, blk_l) * 0.50f + k.ambient;
}

__device__ __forceinline__ float foam_hash(float x, float y) {
    const int32_t qx = (int32_t)x, qy = (int32_t)y;
    return (float)((hash_u32((uint32_t)qx * 374761393u + (uint32_t)qy * 668265263u) >> 8) & 0xFFFFFFu) / 16777216.0f;
}

__device__ inline float foam_mottle(float x, float z, float time) {
    const float sc = (float)(1.0 / 0.7);
    const float p0 = x * sc, p1 = z * sc;
    const float i0 = floorf(p0), i1 = floorf(p1);
    const float w0 = p0 - floorf(p0), w1 = p1 - floorf(p1);
    const float u0 = w0 * w0 * (3.0f - 2.0f * w0), u1 = w1 * w1 * (3.0f - 2.0f * w1);
    const float m0 = r_mix(foam_hash(i0, i1), foam_hash(i0 + 1.0f, i1), u0);
    const float m1 = r_mix(foam_hash(i0, i1 + 1.0f), foam_hash(i0 + 1.0f, i1 + 1.0f), u0);
    const float m = r_mix(m0, m1, u1);
    const float ph = time * FOAM_RATE * 4.0f + foam_hash(i0 * 2.23f, i1 * 2.71f) * 6.28318f;
    return m * (0.5f + 0.5f * r_sin(ph));
}

__device__ __forceinline__ uint32_t face_layer_of(const ResolveK &k, uint32_t id, uint32_t nid) {
    const uint32_t fi = id * 8u + nid;
    return fi < k.n_faces ? (k.faces[fi] & 0xFFFFu) : 0u;
}

__device__ __forceinline__ void voxel_of(uint32_t w, uint32_t v[3]) {
    v[0] = w & 63u;
    v[1] = (w >> 6) & 63u;
    v[2] = (w >> 12) & 63u;
}

// shade_hit at a leg's hit: the sun test of shadow_ray from hit + n * 0.02 vs, then
// shade_hit_legacy without sky specular.
__device__ inline void shade_leg_hit(const ResolveK &k, const WorldHit &w, const float pt[3], const float rd[3],
                                     float dist, float shadow_dist, float out[3]) {
    const FrameK &f = k.f;
    uint32_t v[3];
    voxel_of(w.voxel, v);
    const ChunkK c = ld_chunk(k.p, w.ci);
    const uint32_t id = block_at(k.p, c, v[0], v[1], v[2]);
    float n[3];
    normal_of(w.normal, n);
    float lit = 1.0f;
    if (dot3o(n, f.sun) > 0.0f && f.daylight > 0.01f && (f.flags & VX_FLAG_SHADOWS)) {
        float o[3];
#pragma unroll
        for (int q = 0; q < 3; q++) o[q] = pt[q] + n[q] * (0.02f * c.vs);
        lit = shadow_ray(k.p, f, k.s, o, shadow_dist) ? 0.0f : 1.0f;
    }
    shade_opaque(k, w.ci, v[0], v[1], v[2], id, w.normal, pt, rd, dist, lit, shadow_dist,
                 !(k.bits & RS_FLAT_SECONDARY), false, out);
}

struct WaterCtx {
    WaterNormals nm;
    float body[3], ml, refract_detail, reflect_detail, mirror[3], mirror_t[3];
};

// water_ctx (resolve.wgsl:1078-1167) without the water-look ripples
__device__ inline void water_ctx(const ResolveK &k, uint32_t ci, const uint32_t v[3], uint32_t nid, const float hit[3],
                                 const float rd[3], float t, WaterCtx &o) {
    const FrameK &f = k.f;
    const ChunkK ch = ld_chunk(k.p, ci);
    const float vs = ch.vs;
    float box_min[3], local[3];
#pragma unroll
    for (int q = 0; q < 3; q++) box_min[q] = ch.origin[q] + (float)v[q] * vs;
#pragma unroll
    for (int q = 0; q < 3; q++) local[q] = r_clamp((hit[q] - box_min[q]) / vs, 0.0f, 1.0f);
    const float footprint = fmaxf(t, 1e-3f) * 2.0f * f.tan_half_fov / (float)f.res[1] * 16.0f / vs;
    const float lod = r_clamp(r_log2(fmaxf(footprint, 1e-4f)) + k.mip_bias, 0.0f, 3.0f);
    const uint32_t layer = face_layer_of(k, WATER_ID, nid);
    float texel[4], uv[2];
    if (k.s.bits & SPEC_WATER_REPAIR) {
        uv[0] = uv[1] = 0.5f;
        atlas_sample(k, layer, uv, 3.0f, texel);
    } else {
        face_uv_of(nid, local, uv);
        atlas_sample(k, layer, uv, lod, texel);
    }
    o.ml = medium_light(k, light_at(k.p, ld_chunk_light(k.p, ci), (int32_t)v[0], (int32_t)v[1], (int32_t)v[2]));
#pragma unroll
    for (int q = 0; q < 3; q++) o.body[q] = texel[q] * o.ml;
    o.refract_detail = 1.0f - wgsl_smoothstep(WATER_REFRACT_NEAR, WATER_REFRACT_FAR, t);
    o.reflect_detail = 1.0f - wgsl_smoothstep(WATER_REFLECT_NEAR, WATER_REFLECT_FAR, t);
    water_normals(k.p, f, k.s, nid, hit, rd, t, o.nm);
    wgsl_reflect(rd, o.nm.nw, o.mirror);
    wgsl_reflect(rd, o.nm.nt, o.mirror_t);
    if (o.nm.waved) {
        o.mirror_t[1] = fmaxf(o.mirror_t[1], 0.0f);
        wgsl_normalize(o.mirror_t);
    }
}

// water_refract_leg (resolve.wgsl:1029-1042): rgb and the column depth
__device__ inline void water_refract_leg(const ResolveK &k, const float hit[3], const float n[3], const float rd[3],
                                         float t, float out[4]) {
    const FrameK &f = k.f;
    float rdir[3], ro2[3];
    wgsl_refract(rd, n, 1.0f / WATER_IOR, rdir);
#pragma unroll
    for (int q = 0; q < 3; q++) ro2[q] = hit[q] + rdir[q] * SURFACE_EPS;
    const WorldHit h2 = trace_world<true>(k.p, f, k.s, ro2, rdir, WATER_REFRACT_DIST);
    out[0] = out[1] = out[2] = 0.0f;
    out[3] = WATER_REFRACT_DIST;
    if (h2.hit) {
        float p2[3];
#pragma unroll
        for (int q = 0; q < 3; q++) p2[q] = ro2[q] + rdir[q] * h2.t;
        const float sd = (f.flags & FLAG_WATER_SHADOW_CUT) ? WATER_SHADOW_DIST : WATER_SHADOW_DIST_PRE41;
        shade_leg_hit(k, h2, p2, rdir, t + h2.t, sd, out);
        out[3] = h2.t;
    }
}

// water_reflect_leg (resolve.wgsl:1044-1058): rgb and 1, or all 0 when it hits nothing
__device__ inline void water_reflect_leg(const ResolveK &k, const float hit[3], const float n[3], const float mt[3],
                                         float t, float out[4]) {
    float ro3[3];
#pragma unroll
    for (int q = 0; q < 3; q++) ro3[q] = hit[q] + n[q] * SURFACE_EPS;
    const WorldHit h3 = trace_world<true>(k.p, k.f, k.s, ro3, mt, WATER_REFLECT_DIST);
    out[0] = out[1] = out[2] = out[3] = 0.0f;
    if (!h3.hit) return;
    float p3[3], lit[3], sky[3];
#pragma unroll
    for (int q = 0; q < 3; q++) p3[q] = ro3[q] + mt[q] * h3.t;
    shade_leg_hit(k, h3, p3, mt, t + h3.t, k.f.shadow_dist, lit);
    sky_base(k, mt, 1.0f, sky);
    const float a = 1.0f - transmittance_from(k, hit[1], h3.t, mt);
#pragma unroll
    for (int q = 0; q < 3; q++) out[q] = r_mix(lit[q], sky[q], a);
    out[3] = 1.0f;
}

// water_compose (resolve.wgsl:1169-1207) without the water look
__device__ inline void water_compose(const ResolveK &k, const WaterCtx &x, const float hit[3], const float rd[3],
                                     const float leg_r[4], const float leg_m[4], float out[3]) {
    constexpr float SUN_TINT[3] = {1.0000f, 0.8469f, 0.5647f}, WATER_EXT[3] = {0.280f, 0.055f, 0.020f},
                    FOAM_RGB[3] = {0.92f, 0.97f, 1.0f};
    const FrameK &f = k.f;
    float refr[3], refl[3], sky[3];
#pragma unroll
    for (int q = 0; q < 3; q++) refr[q] = x.body[q];
    float column_depth = WATER_REFRACT_DIST;
    if ((f.flags & FLAG_WATER_REFRACT) && x.refract_detail > 0.0f) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const float tr = wave_exp(-WATER_EXT[q] * k.water_absorb * fmaxf(leg_r[3], 0.0f));
            const float med = leg_r[q] * tr + x.body[q] * (1.0f - tr);
            refr[q] = r_mix(refr[q], med, x.refract_detail);
        }
        column_depth = leg_r[3];
    }
    if ((k.bits & RS_SHORE_FOAM) && x.refract_detail > 0.0f && column_depth < FOAM_DEPTH) {
        const float shall = 1.0f - wgsl_smoothstep(0.0f, FOAM_DEPTH, column_depth);
        const float m = foam_mottle(hit[0], hit[2], f.time * f.wave_speed);
        const float foam = shall * wgsl_smoothstep(FOAM_LO, FOAM_HI, m) * x.refract_detail;
        const float a = r_clamp(foam, 0.0f, 1.0f) * FOAM_MAX;
#pragma unroll
        for (int q = 0; q < 3; q++) refr[q] = r_mix(refr[q], FOAM_RGB[q] * x.ml, a);
    }
    sky_base(k, x.mirror, 1.0f, sky);
    const float g = r_pow(fmaxf(dot3o(x.mirror, f.sun), 0.0f), GLINT_POWER * x.nm.gloss) * GLINT_GAIN * x.nm.gloss *
                    f.daylight;
#pragma unroll
    for (int q = 0; q < 3; q++) refl[q] = sky[q] + SUN_TINT[q] * g;
    if ((f.flags & FLAG_WATER_REFLECT) && x.reflect_detail > 0.0f && leg_m[3] > 0.0f)
#pragma unroll
        for (int q = 0; q < 3; q++) refl[q] = r_mix(refl[q], leg_m[q], x.reflect_detail);
    float skn[3];
    sky_base(k, x.nm.n, 1.0f, skn);
#pragma unroll
    for (int q = 0; q < 3; q++) refl[q] = r_mix(refl[q], skn[q], x.nm.rough * 0.65f);
    const float fr = schlick_water(fmaxf(-dot3o(rd, x.nm.nw), 1e-3f));
#pragma unroll
    for (int q = 0; q < 3; q++) out[q] = r_mix(refr[q], refl[q], fr);
}

// shade_water (resolve.wgsl:1209-1221)
__device__ inline void shade_water(const ResolveK &k, uint32_t ci, const uint32_t v[3], uint32_t nid,
                                   const float hit[3], const float rd[3], float t, float out[3]) {
    WaterCtx x;
    water_ctx(k, ci, v, nid, hit, rd, t, x);
    float leg_r[4] = {0.0f, 0.0f, 0.0f, WATER_REFRACT_DIST}, leg_m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if ((k.f.flags & FLAG_WATER_REFRACT) && x.refract_detail > 0.0f) water_refract_leg(k, hit, x.nm.n, rd, t, leg_r);
    if ((k.f.flags & FLAG_WATER_REFLECT) && x.reflect_detail > 0.0f)
        water_reflect_leg(k, hit, x.nm.n, x.mirror_t, t, leg_m);
    water_compose(k, x, hit, rd, leg_r, leg_m, out);
}

// The decoded primary key: chunk, voxel, micro cell, face.
__device__ __forceinline__ bool key_decode(uint64_t key, uint32_t &ci, uint32_t v[3], uint32_t mc[3], uint32_t &nid) {
    if (key == 0ull) return false;
    nid = (uint32_t)(key >> 37) & 7u;
    ci = (uint32_t)(key >> 24) & 0x1FFFu;
    const uint32_t low = (uint32_t)(key & 0xFFFFFFull);
    v[0] = (low >> 16) & 63u; v[1] = (low >> 8) & 63u; v[2] = low & 63u;
    mc[0] = (low >> 22) & 3u; mc[1] = (low >> 14) & 3u; mc[2] = (low >> 6) & 3u;
    return true;
}

// water_share_safe (resolve.wgsl:1223-1245)
__device__ inline bool water_share_safe(const ResolveK &k, uint32_t px, uint32_t py, uint32_t nid, const float hit[3],
                                        float t) {
    const FrameK &f = k.f;
    if (nid != NORMAL_PY) return false;
    const uint32_t bx = (px >> 1) * WATER_SEC_SPAN, by = (py >> 1) * WATER_SEC_SPAN;
    for (uint32_t y = 0; y < WATER_SEC_SPAN; y++)
        for (uint32_t x = 0; x < WATER_SEC_SPAN; x++) {
            const uint32_t qx = bx + x, qy = by + y;
            if (qx >= f.res[0] || qy >= f.res[1]) continue;
            uint32_t ci, v[3], mc[3], n2;
            if (!key_decode(k.vis[(size_t)qy * f.res[0] + qx], ci, v, mc, n2)) return false;
            if (n2 != nid || ci >= k.p.n_chunks) return false;
            const ChunkK c = ld_chunk(k.p, ci);
            if (block_at(k.p, c, v[0], v[1], v[2]) != WATER_ID) return false;
            float rd[3];
            ray_dir(f, (float)qx + 0.5f, (float)qy + 0.5f, rd[0], rd[1], rd[2]);
            const float td = hit_t(f, c, v[0], v[1], v[2], mc[0], mc[1], mc[2], n2, rd);
            const float hy = f.cam[1] + rd[1] * td;
            if (fabsf(hy - hit[1]) > 0.01f || fabsf(td - t) > fmaxf(0.1f, fminf(1.0f, t * 0.02f))) return false;
        }
    return true;
}

__device__ __forceinline__ float h16f(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }
__device__ __forceinline__ uint16_t f16h(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }

// shade_water_sec (resolve.wgsl:1247-1259)
__device__ inline void shade_water_sec(const ResolveK &k, uint32_t ci, const uint32_t v[3], uint32_t nid,
                                       const float hit[3], const float rd[3], float t, uint32_t px, uint32_t py,
                                       float out[3]) {
    if ((k.s.bits & SPEC_WATER_REPAIR) && !water_share_safe(k, px, py, nid, hit, t)) {
        shade_water(k, ci, v, nid, hit, rd, t, out);
        return;
    }
    WaterCtx x;
    water_ctx(k, ci, v, nid, hit, rd, t, x);
    const size_t o = (size_t)(py >> 1) * k.hw + (px >> 1);
    const ushort4 r = k.sec_r[o], m = k.sec_m[o];
    const float leg_r[4] = {h16f(r.x), h16f(r.y), h16f(r.z), h16f(r.w)};
    const float leg_m[4] = {h16f(m.x), h16f(m.y), h16f(m.z), h16f(m.w)};
    water_compose(k, x, hit, rd, leg_r, leg_m, out);
}

// One leg of shade_glass from origin along dir: the hit shaded (shade_water on water) and
// mixed toward the sky by the fog over the leg; out keeps its value when it hits nothing.
__device__ inline void glass_leg(const ResolveK &k, const float hit[3], const float origin[3], const float dir[3],
                                 float t, float out[3]) {
    const WorldHit h = trace_world<false>(k.p, k.f, k.s, origin, dir, fmaxf(GLASS_TRANSMIT_DIST, k.f.far_));
    if (!h.hit) return;
    float p2[3], col[3], sky[3];
#pragma unroll
    for (int q = 0; q < 3; q++) p2[q] = origin[q] + dir[q] * h.t;
    uint32_t v[3];
    voxel_of(h.voxel, v);
    const ChunkK c = ld_chunk(k.p, h.ci);
    if (block_at(k.p, c, v[0], v[1], v[2]) == WATER_ID) shade_water(k, h.ci, v, h.normal, p2, dir, t + h.t, col);
    else shade_leg_hit(k, h, p2, dir, t + h.t, GLASS_SHADOW_DIST, col);
    sky_base(k, dir, 1.0f, sky);
    const float a = 1.0f - transmittance_from(k, hit[1], h.t, dir);
#pragma unroll
    for (int q = 0; q < 3; q++) out[q] = r_mix(col[q], sky[q], a);
}

// shade_glass (resolve.wgsl:1261-1319)
__device__ inline void shade_glass(const ResolveK &k, uint32_t ci, const uint32_t v[3], uint32_t nid,
                                   const float hit[3], const float rd[3], float t, float out[3]) {
    const FrameK &f = k.f;
    const ChunkK ch = ld_chunk(k.p, ci);
    const float vs = ch.vs;
    float n[3];
    normal_of(nid, n);
    if (dot3o(n, rd) > 0.0f)
#pragma unroll
        for (int q = 0; q < 3; q++) n[q] = -n[q];
    float box_min[3], local[3], uv[2], texel[4];
#pragma unroll
    for (int q = 0; q < 3; q++) box_min[q] = ch.origin[q] + (float)v[q] * vs;
#pragma unroll
    for (int q = 0; q < 3; q++) local[q] = r_clamp((hit[q] - box_min[q]) / vs, 0.0f, 1.0f);
    const float footprint = fmaxf(t, 1e-3f) * 2.0f * f.tan_half_fov / (float)f.res[1] * 16.0f / vs;
    const float lod = r_clamp(r_log2(fmaxf(footprint, 1e-4f)) + k.mip_bias, 0.0f, 3.0f);
    face_uv_of(nid, local, uv);
    atlas_sample(k, face_layer_of(k, GLASS_ID, nid), uv, lod, texel);
    float ro2[3], through[3], mirror[3], refl[3];
#pragma unroll
    for (int q = 0; q < 3; q++) ro2[q] = hit[q] + rd[q] * SURFACE_EPS;
    sky_base(k, rd, 1.0f, through);
    glass_leg(k, hit, ro2, rd, t, through);
    wgsl_reflect(rd, n, mirror);
    sky_base(k, mirror, 1.0f, refl);
    if (k.s.bits & SPEC_GLASS_REFLECT) {
        float origin[3];
#pragma unroll
        for (int q = 0; q < 3; q++) origin[q] = hit[q] + n[q] * SURFACE_EPS;
        glass_leg(k, hit, origin, mirror, t, refl);
    }
    const float fr = schlick_glass(fmaxf(-dot3o(rd, n), 1e-3f));
#pragma unroll
    for (int q = 0; q < 3; q++) out[q] = r_mix(through[q] * texel[q], refl[q], fr);
}

// water_sec (resolve.wgsl:1469-1535): one thread per 2x2 block; the legs from the block's
// water pixel of least hit distance (the first in x-major order on ties), stored as f16.
__global__ __launch_bounds__(64) void k_water_sec(ResolveK k) {
    const FrameK &f = k.f;
    const uint32_t bxs = (k.hw + 7u) / 8u;
    const uint32_t hx = (blockIdx.x % bxs) * 8u + (threadIdx.x & 7u), hy = (blockIdx.x / bxs) * 8u + (threadIdx.x >> 3);
    if (hx >= k.hw || hy >= k.hh) return;
    float leg_r[4] = {0.0f, 0.0f, 0.0f, WATER_REFRACT_DIST}, leg_m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float best_t = 3.4e38f;
    bool has = false;
    uint32_t bci = 0, bv[3] = {0, 0, 0}, bnid = 0, bpx = 0, bpy = 0;
    for (uint32_t i = 0; i < WATER_SEC_SPAN; i++)
        for (uint32_t j = 0; j < WATER_SEC_SPAN; j++) {
            const uint32_t px = WATER_SEC_SPAN * hx + i, py = WATER_SEC_SPAN * hy + j;
            if (px >= f.res[0] || py >= f.res[1]) continue;
            uint32_t ci, v[3], mc[3], nid;
            if (!key_decode(k.vis[(size_t)py * f.res[0] + px], ci, v, mc, nid) || ci >= k.p.n_chunks) continue;
            const ChunkK c = ld_chunk(k.p, ci);
            if (block_at(k.p, c, v[0], v[1], v[2]) != WATER_ID) continue;
            float rd0[3];
            ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd0[0], rd0[1], rd0[2]);
            const float ts = hit_t(f, c, v[0], v[1], v[2], mc[0], mc[1], mc[2], nid, rd0);
            if (ts < best_t) {
                best_t = ts; bci = ci; bnid = nid; bpx = px; bpy = py; has = true;
                bv[0] = v[0]; bv[1] = v[1]; bv[2] = v[2];
            }
        }
    if (has) {
        float rd[3], hit[3];
        ray_dir(f, (float)bpx + 0.5f, (float)bpy + 0.5f, rd[0], rd[1], rd[2]);
#pragma unroll
        for (int q = 0; q < 3; q++) hit[q] = f.cam[q] + rd[q] * best_t;
        WaterCtx x;
        water_ctx(k, bci, bv, bnid, hit, rd, best_t, x);
        if ((f.flags & FLAG_WATER_REFRACT) && x.refract_detail > 0.0f) water_refract_leg(k, hit, x.nm.n, rd, best_t, leg_r);
        if ((f.flags & FLAG_WATER_REFLECT) && x.reflect_detail > 0.0f)
            water_reflect_leg(k, hit, x.nm.n, x.mirror_t, best_t, leg_m);
    }
    const size_t o = (size_t)hy * k.hw + hx;
    k.sec_r[o] = make_ushort4(f16h(leg_r[0]), f16h(leg_r[1]), f16h(leg_r[2]), f16h(leg_r[3]));
    k.sec_m[o] = make_ushort4(f16h(leg_m[0]), f16h(leg_m[1]), f16h(leg_m[2]), f16h(leg_m[3]));
}

// The glass and water pixels of the resolve (lanes 1 and 2), one wave per 8x8 tile; other
// pixels are lane 0's.
__global__ __launch_bounds__(64) void k_resolve_lanes(ResolveK k) {
    const FrameK &f = k.f;
    const uint32_t tile = blockIdx.x;
    if (tile >= f.tiles[0] * f.tiles[1]) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t px = (tile % f.tiles[0]) * 8u + (lane & 7u), py = (tile / f.tiles[0]) * 8u + (lane >> 3);
    if (px >= f.res[0] || py >= f.res[1]) return;
    const size_t pix = (size_t)py * f.res[0] + px;
    const uint64_t key = k.vis[pix];
    if (key == 0ull) return;
    const KeyHit h = key_hit(k.p, f, key, px, py);
    if (!h.ok) return;
    const bool glass = (k.bits & RS_GLASS) && h.id == GLASS_ID;
    if (h.id != WATER_ID && !glass) return;
    const uint32_t low = (uint32_t)(key & 0xFFFFFFull);
    const uint32_t ci = (uint32_t)(key >> 24) & 0x1FFFu;
    const uint32_t v[3] = {(low >> 16) & 63u, (low >> 8) & 63u, low & 63u};
    float color[3];
    if (glass) shade_glass(k, ci, v, h.nid, h.hit, h.rd, h.t, color);
    else if (k.bits & RS_WATER_SEC) shade_water_sec(k, ci, v, h.nid, h.hit, h.rd, h.t, px, py, color);
    else shade_water(k, ci, v, h.nid, h.hit, h.rd, h.t, color);
    const float alpha = 1.0f - transmittance_from(k, f.cam[1] + 0.0f * h.rd[1], h.t - 0.0f, h.rd);
    float sky[3];
    sky_base(k, h.rd, 1.0f, sky);
#pragma unroll
    for (int q = 0; q < 3; q++) color[q] = r_mix(color[q], sky[q], alpha);
    k.out[pix] = make_float4(color[0], color[1], color[2], 1.0f);
}

// heat (resolve.wgsl:36-44)
__device__ __forceinline__ void heat_rgb(float v, float o[3]) {
    const float t = r_clamp(v, 0.0f, 1.0f);
    const float c[3] = {wgsl_smoothstep(0.35f, 0.85f, t),
                        wgsl_smoothstep(0.0f, 0.45f, t) - wgsl_smoothstep(0.65f, 1.0f, t),
                        1.0f - wgsl_smoothstep(0.0f, 0.45f, t)};
#pragma unroll
    for (int q = 0; q < 3; q++) o[q] = r_pow(c[q], 2.2f);
}

__device__ __forceinline__ float f16_word(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

__global__ __launch_bounds__(64) void k_resolve(ResolveK k) {
    const FrameK &f = k.f;
    const uint32_t tile = blockIdx.x;
    if (tile >= f.tiles[0] * f.tiles[1]) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t px = (tile % f.tiles[0]) * 8u + (lane & 7u), py = (tile / f.tiles[0]) * 8u + (lane >> 3);
    if (px >= f.res[0] || py >= f.res[1]) return;
    const size_t pix = (size_t)py * f.res[0] + px;
    float rd[3], color[3];
    ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd[0], rd[1], rd[2]);
    if (f.flags & VX_FLAG_HEATMAP) {
        heat_rgb((float)(k.heat ? k.heat[pix] : 0u) / 300.0f, color);
        k.out[pix] = make_float4(color[0], color[1], color[2], 1.0f);
        return;
    }
    const uint64_t key = k.vis[pix];
    if (key == 0ull) {
        sky_base(k, rd, 1.0f, color);
    } else {
        const KeyHit h = key_hit(k.p, f, key, px, py);
        if (!h.ok || h.id == WATER_ID || ((k.bits & RS_GLASS) && h.id == GLASS_ID)) {
            k.out[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            return;
        }
        const uint32_t low = (uint32_t)(key & 0xFFFFFFull);
        const uint32_t ci = (uint32_t)(key >> 24) & 0x1FFFu;
        // the sun visibility of shadow_horizontal / shadow_vertical without soft shadows
        // (shadow.wgsl:61-116): clamp(1 - occ, 0, 1), occ = blocker distance > 0
        const float d = f16_word(k.sun ? k.sun[pix] : (uint16_t)0);
        const float occ = d > 0.0f ? 1.0f : 0.0f;
        const float lit = k.soft_mask ? (float)k.soft_mask[pix] / 255.0f : r_clamp(1.0f - occ, 0.0f, 1.0f);
        shade_opaque(k, ci, (low >> 16) & 63u, (low >> 8) & 63u, low & 63u, h.id, h.nid, h.hit, h.rd, h.t, lit,
                     f.shadow_dist, true, true, color);
        const float alpha = 1.0f - transmittance_from(k, f.cam[1] + 0.0f * h.rd[1], h.t - 0.0f, h.rd);
        float sky[3];
        sky_base(k, h.rd, 1.0f, sky);
#pragma unroll
        for (int q = 0; q < 3; q++) color[q] = r_mix(color[q], sky[q], alpha);
    }
    k.out[pix] = make_float4(color[0], color[1], color[2], 1.0f);
}

// ---- present (blit.wgsl:18-93), as the oracle's vxo_present ----
constexpr float KNEE = 0.75f;
__device__ __forceinline__ float tone_knee(float x) {
    const float s = (1.0f - KNEE) * (1.0f - KNEE);
    const float shoulder = 1.0f - s / (fmaxf(x, KNEE) + (1.0f - 2.0f * KNEE));
    return x < KNEE ? x : shoulder;
}
__device__ __forceinline__ float aces_fit(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return r_clamp((x * (a * x + b)) / (x * (c * x + d) + e), 0.0f, 1.0f);
}
__device__ __forceinline__ float linear_to_srgb(float c) {
    const float lo = c * 12.92f;
    const float hi = 1.055f * r_pow(fmaxf(c, 0.0031308f), 1.0f / 2.4f) - 0.055f;
    return c <= 0.0031308f ? lo : hi;
}
__device__ __forceinline__ uint32_t unorm8(float c) { return (uint32_t)floorf(r_clamp(c, 0.0f, 1.0f) * 255.0f + 0.5f); }

__global__ __launch_bounds__(256) void k_present(const float4 *__restrict__ src, uint32_t n, uint32_t encode,
                                                 uint32_t tone, uint32_t grade, float strength,
                                                 uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 v = src[i];
    const float hdr[3] = {fmaxf(v.x, 0.0f), fmaxf(v.y, 0.0f), fmaxf(v.z, 0.0f)};
    float c[3];
#pragma unroll
    for (int k = 0; k < 3; k++) c[k] = tone == 1u ? aces_fit(hdr[k]) : tone_knee(hdr[k]);
    const float st = r_clamp(strength, 0.0f, 1.0f);
    if (grade == 1u) {
        const float before[3] = {c[0], c[1], c[2]};
        const float luma = c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f;
        const float bal[3] = {r_clamp(c[0] * 1.05f, 0.0f, 1.0f), r_clamp(c[1] * 1.0f, 0.0f, 1.0f),
                              r_clamp(c[2] * 0.93f, 0.0f, 1.0f)};
#pragma unroll
        for (int k = 0; k < 3; k++) c[k] = r_clamp(luma + (bal[k] - luma) * 1.06f, 0.0f, 1.0f);
        if (strength < 1.0f) {
#pragma unroll
            for (int k = 0; k < 3; k++) c[k] = r_mix(before[k], c[k], st);
        }
    }
    if (grade == 2u) {
        const float before[3] = {c[0], c[1], c[2]};
        const float luma = c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f;
        float g[3];
#pragma unroll
        for (int k = 0; k < 3; k++) g[k] = r_mix(c[k], luma, 0.12f);
#pragma unroll
        for (int k = 0; k < 3; k++) g[k] = r_mix(g[k], g[k] * g[k] * (3.0f - 2.0f * g[k]), 0.45f);
        const float shad = r_pow(r_clamp(1.0f - luma * 2.2f, 0.0f, 1.0f), 2.0f);
        const float hi = r_pow(r_clamp(luma, 0.0f, 1.0f), 3.0f);
        const float m[3] = {1.0f - 0.05f * shad + 0.03f * hi, 1.0f - 0.02f * shad + 0.02f * hi,
                            1.0f + 0.07f * shad - 0.01f * hi};
#pragma unroll
        for (int k = 0; k < 3; k++) c[k] = r_clamp(g[k] * m[k], 0.0f, 1.0f);
        if (strength < 1.0f) {
#pragma unroll
            for (int k = 0; k < 3; k++) c[k] = r_mix(before[k], c[k], st);
        }
    }
    if (encode != 0u) {
#pragma unroll
        for (int k = 0; k < 3; k++) c[k] = linear_to_srgb(c[k]);
    }
    out[i] = unorm8(c[0]) | (unorm8(c[1]) << 8) | (unorm8(c[2]) << 16) | (255u << 24);
}

// ---- TAA (taa.wgsl:1-127), as the oracle's vxo_taa ----
// One wave per 8x8 tile. The 10x10 colour neighbourhood (edge texels repeated) is staged
// in LDS as the trace output holds it (the resolve's f32 colour rounded to f16); each
// lane reprojects its pixel through prev_view_proj, filters the f16 history at that
// position (Catmull-Rom in five clamp-to-edge bilinear taps), clips it in YCoCg to the
// neighbourhood's mean +- 1 sigma and mixes it with the colour by taa_feedback. The
// output is stored as f16 RGBA (alpha 1) and is the next frame's history.
struct TaaK {
    Pools p;
    FrameK f;
    float prev[16];  // prev_view_proj, column major
    float feedback;
    const unsigned long long *__restrict__ vis;
    const float4 *__restrict__ src;
    const ushort4 *__restrict__ hist;
    ushort4 *__restrict__ out;
};

__device__ __forceinline__ float f16r(float x) { return (float)(_Float16)x; }
__device__ __forceinline__ float h2f(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }
__device__ __forceinline__ uint16_t f2h(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }

__device__ __forceinline__ void rgb_to_ycocg(const float c[3], float o[3]) {
    o[0] = 0.25f * c[0] + 0.5f * c[1] + 0.25f * c[2];
    o[1] = 0.5f * c[0] - 0.5f * c[2];
    o[2] = -0.25f * c[0] + 0.5f * c[1] - 0.25f * c[2];
}
__device__ __forceinline__ void ycocg_to_rgb(const float c[3], float o[3]) {
    const float t = c[0] - c[2];
    o[0] = t + c[1];
    o[1] = c[0] + c[2];
    o[2] = t - c[1];
}

__device__ __forceinline__ void hist_bilinear(const ushort4 *hist, uint32_t w, uint32_t h, float u, float v,
                                              float o[3]) {
    const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const int32_t x0 = clamp_texel((int32_t)fx, (int32_t)w), x1 = clamp_texel((int32_t)fx + 1, (int32_t)w);
    const int32_t y0 = clamp_texel((int32_t)fy, (int32_t)h), y1 = clamp_texel((int32_t)fy + 1, (int32_t)h);
    const ushort4 t00 = hist[(size_t)y0 * w + x0], t10 = hist[(size_t)y0 * w + x1];
    const ushort4 t01 = hist[(size_t)y1 * w + x0], t11 = hist[(size_t)y1 * w + x1];
    const float c00[3] = {h2f(t00.x), h2f(t00.y), h2f(t00.z)}, c10[3] = {h2f(t10.x), h2f(t10.y), h2f(t10.z)};
    const float c01[3] = {h2f(t01.x), h2f(t01.y), h2f(t01.z)}, c11[3] = {h2f(t11.x), h2f(t11.y), h2f(t11.z)};
#pragma unroll
    for (int k = 0; k < 3; k++) o[k] = r_mix(r_mix(c00[k], c10[k], a), r_mix(c01[k], c11[k], a), b);
}

__global__ __launch_bounds__(64) void k_taa(TaaK k) {
    __shared__ float tile[100][3];
    const FrameK &f = k.f;
    const uint32_t w = f.res[0], h = f.res[1];
    const uint32_t t = blockIdx.x;
    if (t >= f.tiles[0] * f.tiles[1]) return;
    const int32_t gx = (int32_t)(t % f.tiles[0]), gy = (int32_t)(t / f.tiles[0]);
    const uint32_t lane = threadIdx.x;
    for (uint32_t q = lane; q < 100u; q += 64u) {
        const int32_t x = min(max(gx * 8 - 1 + (int32_t)(q % 10u), 0), (int32_t)w - 1);
        const int32_t y = min(max(gy * 8 - 1 + (int32_t)(q / 10u), 0), (int32_t)h - 1);
        const float4 c = k.src[(size_t)y * w + x];
        tile[q][0] = f16r(c.x);
        tile[q][1] = f16r(c.y);
        tile[q][2] = f16r(c.z);
    }
    __syncthreads();
    const uint32_t lx = lane & 7u, ly = lane >> 3;
    const uint32_t px = (uint32_t)gx * 8u + lx, py = (uint32_t)gy * 8u + ly;
    if (px >= w || py >= h) return;
    float cur[3], o[3];
#pragma unroll
    for (int q = 0; q < 3; q++) o[q] = cur[q] = tile[(ly + 1u) * 10u + lx + 1u][q];
    const bool heat = (f.flags & VX_FLAG_HEATMAP) != 0u;
    if (k.feedback < 1.0f && !heat) {
        float rd[3];
        ray_dir(f, (float)px + 0.5f, (float)py + 0.5f, rd[0], rd[1], rd[2]);
        const uint64_t key = k.vis[(size_t)py * w + px];
        float q4[4];
        if (key == 0ull) {
            q4[0] = rd[0]; q4[1] = rd[1]; q4[2] = rd[2]; q4[3] = 0.0f;
        } else {
            const uint32_t ci = (uint32_t)(key >> 24) & 0x1FFFu, low = (uint32_t)(key & 0xFFFFFFull);
            const ChunkK c = ld_chunk(k.p, ci);
            const float th = hit_t(f, c, (low >> 16) & 63u, (low >> 8) & 63u, low & 63u, (low >> 22) & 3u,
                                   (low >> 14) & 3u, (low >> 6) & 3u, (uint32_t)(key >> 37) & 7u, rd);
#pragma unroll
            for (int q = 0; q < 3; q++) q4[q] = f.cam[q] + rd[q] * th;
            q4[3] = 1.0f;
        }
        float clip[4];
#pragma unroll
        for (int r = 0; r < 4; r++)
            clip[r] = ((k.prev[r] * q4[0] + k.prev[4 + r] * q4[1]) + k.prev[8 + r] * q4[2]) + k.prev[12 + r] * q4[3];
        bool ok = clip[3] > 1e-6f;
        float pr[2] = {0.0f, 0.0f};
        if (ok) {
            const float nx = clip[0] / clip[3], ny = clip[1] / clip[3];
            const float rx = (float)w, ry = (float)h;
            pr[0] = (nx * 0.5f + 0.5f) * rx - f.jitter[0];
            pr[1] = (0.5f - ny * 0.5f) * ry - f.jitter[1];
            ok = !(pr[0] < 0.0f || pr[1] < 0.0f || pr[0] > rx || pr[1] > ry);
        }
        if (ok) {
            float m1[3] = {0.0f, 0.0f, 0.0f}, m2[3] = {0.0f, 0.0f, 0.0f};
            for (uint32_t i = 0; i < 9u; i++) {
                float y[3];
                rgb_to_ycocg(tile[(ly + i / 3u) * 10u + lx + i % 3u], y);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    m1[q] = m1[q] + y[q];
                    m2[q] = m2[q] + y[q] * y[q];
                }
            }
            // history_bicubic (taa.wgsl:28-58)
            const float res[2] = {(float)w, (float)h};
            float wt[5], tap[5][2], w0[2], w3[2], w12[2], t0[2], t3[2], t12[2];
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const float tex1 = floorf(pr[a] - 0.5f) + 0.5f;
                const float fr = pr[a] - tex1, f2 = fr * fr, f3 = f2 * fr;
                w0[a] = -0.5f * f3 + f2 - 0.5f * fr;
                const float w1 = 1.5f * f3 - 2.5f * f2 + 1.0f;
                const float w2 = -1.5f * f3 + 2.0f * f2 + 0.5f * fr;
                w3[a] = 0.5f * f3 - 0.5f * f2;
                w12[a] = w1 + w2;
                t0[a] = (tex1 - 1.0f) / res[a];
                t3[a] = (tex1 + 2.0f) / res[a];
                t12[a] = (tex1 + w2 / w12[a]) / res[a];
            }
            tap[0][0] = t12[0]; tap[0][1] = t0[1];  wt[0] = w12[0] * w0[1];
            tap[1][0] = t0[0];  tap[1][1] = t12[1]; wt[1] = w0[0] * w12[1];
            tap[2][0] = t12[0]; tap[2][1] = t12[1]; wt[2] = w12[0] * w12[1];
            tap[3][0] = t3[0];  tap[3][1] = t12[1]; wt[3] = w3[0] * w12[1];
            tap[4][0] = t12[0]; tap[4][1] = t3[1];  wt[4] = w12[0] * w3[1];
            float hb[3] = {0.0f, 0.0f, 0.0f};
            for (int i = 0; i < 5; i++) {
                float sm[3];
                hist_bilinear(k.hist, w, h, tap[i][0], tap[i][1], sm);
#pragma unroll
                for (int q = 0; q < 3; q++) hb[q] = hb[q] + sm[q] * wt[i];
            }
            const float sum = wt[0] + wt[1] + wt[2] + wt[3] + wt[4];
#pragma unroll
            for (int q = 0; q < 3; q++) hb[q] = fmaxf(hb[q] / sum, 0.0f);
            float hy[3], centre[3], v[3], n[3], cl[3], back[3];
            rgb_to_ycocg(hb, hy);
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const float mean = m1[q] / 9.0f;
                const float sigma = sqrtf(fmaxf(m2[q] / 9.0f - mean * mean, 0.0f));
                const float lo = mean - 1.0f * sigma, hi = mean + 1.0f * sigma;
                centre[q] = 0.5f * (hi + lo);
                const float extent = fmaxf(0.5f * (hi - lo), 1e-5f);
                v[q] = hy[q] - centre[q];
                n[q] = fabsf(v[q] / extent);
            }
            const float m = fmaxf(n[0], fmaxf(n[1], n[2]));
#pragma unroll
            for (int q = 0; q < 3; q++) cl[q] = m > 1.0f ? centre[q] + v[q] / m : hy[q];
            ycocg_to_rgb(cl, back);
#pragma unroll
            for (int q = 0; q < 3; q++) o[q] = r_mix(fmaxf(back[q], 0.0f), cur[q], k.feedback);
        }
    }
    k.out[(size_t)py * w + px] = make_ushort4(f2h(o[0]), f2h(o[1]), f2h(o[2]), (uint16_t)0x3C00u);
}

// The present of a TAA output: the f16 history read as the blit samples it.
__global__ __launch_bounds__(256) void k_half_to_float(const ushort4 *__restrict__ src, uint32_t n,
                                                       float4 *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ushort4 v = src[i];
    out[i] = make_float4(h2f(v.x), h2f(v.y), h2f(v.z), h2f(v.w));
}

}  // namespace vxd

using namespace vxd;

#define HIPX VX_HIPX

namespace {

// The shaft buffer (three slices), zeroed when first made, as the reference's buffer starts.
int shaft_buffer(vx_ctx *c) {
    if (c->shaft.p) return VX_OK;
    int rc;
    if ((rc = vxh::ensure(c->shaft, (size_t)3 * SHAFT_SLICE * 4))) return rc;
    return vxh::fill_u32(c->stream, c->shaft.p, 0u, (size_t)3 * SHAFT_SLICE);
}

// ---- soft shadows: shadow_horizontal / shadow_vertical (shadow.wgsl:49-116) ----
constexpr float SUN_ANGLE = 0.0093f, SOFT_PENUMBRA = 1.5f;
constexpr int32_t SHADOW_RADIUS_LIMIT = 4;

struct SGuide {
    float hit[3], n[3], t, scale;
    bool valid;
};

// shadow_guide (shadow.wgsl:8-30) at pixel (x, y)
__device__ inline SGuide soft_guide(const ResolveK &k, int32_t x, int32_t y) {
    SGuide g;
    g.valid = false;
    const FrameK &f = k.f;
    if (x < 0 || y < 0 || x >= (int32_t)f.res[0] || y >= (int32_t)f.res[1]) return g;
    const uint64_t key = k.vis[(size_t)y * f.res[0] + (uint32_t)x];
    const KeyHit h = key_hit(k.p, f, key, (uint32_t)x, (uint32_t)y);
    if (!h.ok || h.id == WATER_ID || ((k.bits & RS_GLASS) && h.id == GLASS_ID)) return g;
    normal_of(h.nid, g.n);
    if (h.nid >= NORMAL_CROSS_A && dot3o(g.n, h.rd) > 0.0f) {
        g.n[0] = -g.n[0]; g.n[1] = -g.n[1]; g.n[2] = -g.n[2];
    }
    if (!(dot3o(g.n, f.sun) > 0.0f)) return g;
    g.t = h.t;
#pragma unroll
    for (int q = 0; q < 3; q++) g.hit[q] = h.hit[q];
    g.scale = ld_chunk(k.p, (uint32_t)(key >> 24) & 0x1FFFu).vs;
    g.valid = g.t >= 0.0f;
    return g;
}

__device__ __forceinline__ float shadow_radius(const FrameK &f, float distance, float depth) {
    const float proj = (float)f.res[1] / (2.0f * fmaxf(f.tan_half_fov, 0.001f));
    const float r = distance * SUN_ANGLE * SOFT_PENUMBRA * proj / fmaxf(depth, 0.01f);
    return r >= 0.5f ? fminf(r, (float)SHADOW_RADIUS_LIMIT) : 0.0f;
}

__device__ __forceinline__ bool shadow_compatible(const SGuide &a, const SGuide &b) {
    if (!b.valid || dot3o(a.n, b.n) < 0.95f) return false;
    const float tol = fmaxf(0.01f, 0.01f * fminf(a.scale, b.scale));
    const float d[3] = {b.hit[0] - a.hit[0], b.hit[1] - a.hit[1], b.hit[2] - a.hit[2]};
    return fabsf(dot3o(d, a.n)) <= tol && fabsf(dot3o(d, b.n)) <= tol;
}

__device__ __forceinline__ bool soft_active(const FrameK &f) {
    return (f.flags & VX_FLAG_SHADOWS) && !(f.flags & VX_FLAG_HEATMAP) && f.daylight > 0.01f;
}

// shadow_horizontal: per pixel the occlusion averaged along its row over compatible
// neighbours within their penumbra radius, and the largest radius met (RG16F).
__global__ __launch_bounds__(256) void k_soft_h(ResolveK k) {
    const uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    const FrameK &f = k.f;
    if (pix >= f.res[0] * f.res[1]) return;
    const int32_t x = (int32_t)(pix % f.res[0]), y = (int32_t)(pix / f.res[0]);
    const float d = f16_word(k.sun[pix]);
    const float occ = d > 0.0f ? 1.0f : 0.0f;
    float rx = occ, ry = 0.0f;
    if (soft_active(f)) {
        const SGuide center = soft_guide(k, x, y);
        if (center.valid) {
            const float rc = shadow_radius(f, d, center.t);
            float sum = occ, weight = 1.0f, support = rc;
            for (int32_t q = -SHADOW_RADIUS_LIMIT; q <= SHADOW_RADIUS_LIMIT; q++) {
                if (q == 0) continue;
                const SGuide g = soft_guide(k, x + q, y);
                if (!shadow_compatible(center, g)) continue;
                const float nd = f16_word(k.sun[pix + q]);
                const float r = fmaxf(rc, shadow_radius(f, nd, g.t));
                if (r < fabsf((float)q)) continue;
                const float w = wave_exp(-2.0f * (float)(q * q) / fmaxf(r * r, 0.25f));
                sum = sum + (nd > 0.0f ? 1.0f : 0.0f) * w;
                weight = weight + w;
                support = fmaxf(support, r);
            }
            rx = sum / weight;
            ry = support;
        }
    }
    k.soft_h[pix] = make_ushort2(f16h(rx), f16h(ry));
}

// shadow_vertical: the row averages averaged along the column likewise; the mask is
// clamp(1 - occ, 0, 1) stored as R8 unorm.
__global__ __launch_bounds__(256) void k_soft_v(ResolveK k) {
    const uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    const FrameK &f = k.f;
    if (pix >= f.res[0] * f.res[1]) return;
    const int32_t x = (int32_t)(pix % f.res[0]), y = (int32_t)(pix / f.res[0]);
    const ushort2 h = k.soft_h[pix];
    const float hx = f16_word(h.x), hy = f16_word(h.y);
    float occ = hx;
    if (soft_active(f)) {
        const SGuide center = soft_guide(k, x, y);
        if (center.valid) {
            float sum = hx, weight = 1.0f;
            for (int32_t q = -SHADOW_RADIUS_LIMIT; q <= SHADOW_RADIUS_LIMIT; q++) {
                if (q == 0) continue;
                const SGuide g = soft_guide(k, x, y + q);
                if (!shadow_compatible(center, g)) continue;
                const ushort2 o = k.soft_h[(size_t)(y + q) * f.res[0] + (uint32_t)x];
                const float r = fmaxf(hy, f16_word(o.y));
                if (r < fabsf((float)q)) continue;
                const float w = wave_exp(-2.0f * (float)(q * q) / fmaxf(r * r, 0.25f));
                sum = sum + f16_word(o.x) * w;
                weight = weight + w;
            }
            occ = sum / weight;
        }
    }
    k.soft_mask[pix] = (uint8_t)floorf(r_clamp(1.0f - occ, 0.0f, 1.0f) * 255.0f + 0.5f);
}

// make_probe's fill (src/render/mod.rs:2270-2287): value, or the hashed noise.
__global__ __launch_bounds__(256) void k_probe_fill(ushort4 *__restrict__ t, float value, int noise) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= PROBE_TEXELS) return;
    uint16_t h4[4];
#pragma unroll
    for (uint32_t c = 0; c < 4u; c++) {
        float v = value;
        if (noise) {
            const uint32_t h = (uint32_t)i * 2654435761u + c * 0x9E3779B9u;
            v = 0.5f + (float)(h >> 8) / (float)(1u << 25);
        }
        h4[c] = __builtin_bit_cast(uint16_t, (_Float16)v);
    }
    t[i] = make_ushort4(h4[0], h4[1], h4[2], h4[3]);
}

// write_probe (src/render/mod.rs:1398-1445): one chunk's 6 x 16^3 probes into its block of
// each slab, bounce rgb and occlusion as f16.
__global__ __launch_bounds__(256) void k_probe_write(ushort4 *__restrict__ t, const float *__restrict__ occl,
                                                     const float *__restrict__ bounce, uint32_t bx, uint32_t by,
                                                     uint32_t bz) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;  // axis * 4096 + probe
    if (j >= 6u * 4096u) return;
    const uint32_t axis = j >> 12, i = j & 4095u;
    const uint32_t px = i & 15u, py = (i >> 4) & 15u, pz = i >> 8;
    const uint32_t x = bx + px, y = axis * PROBE_Y + by + py, z = bz + pz;
    auto h = [](float v) { return __builtin_bit_cast(uint16_t, (_Float16)v); };  // round to nearest even
    t[((size_t)z * PROBE_H + y) * PROBE_XZ + x] =
        make_ushort4(h(bounce[j * 3u]), h(bounce[j * 3u + 1u]), h(bounce[j * 3u + 2u]), h(occl[j]));
}

// The probe field, made filled with ones as the default make_probe makes it.
int probe_buffer(vx_ctx *c) {
    if (c->probe.p) return VX_OK;
    int rc;
    if ((rc = vxh::ensure(c->probe, PROBE_TEXELS * 8))) return rc;
    k_probe_fill<<<(unsigned)((PROBE_TEXELS + 255) / 256), 256, 0, c->stream>>>((ushort4 *)c->probe.p, 1.0f, 0);
    VX_LAUNCHED();
    return VX_OK;
}

// The resolve switches of a pipeline key, or VX_E_UNSUPPORTED for what is not restated.
int resolve_bits(uint32_t spec, uint32_t spec_hi, const vx_gpu_frame &g, uint32_t &bits) {
    if (g.flags & VX_FLAG_HEATMAP) {
        bits = 0;
        return VX_OK;
    }
    if ((g.flags & VX_FLAG_UNDERWATER) || (spec & VX_FLAG_FOLIAGE)) return VX_E_UNSUPPORTED;
    if (spec_hi & (8192u | 16384u)) return VX_E_UNSUPPORTED;  // foliage-rich and canopy-relief albedo
    if (spec_hi & 4096u) return VX_E_UNSUPPORTED;                           // the water look (ripples, murk)
    if (g.cloud_cover > 0.0f) return VX_E_UNSUPPORTED;                      // cloud deck
    const bool terrain_shafts = (spec & 8192u) && (g.flags & 8192u);
    if (g.godray_strength > 0.0f && terrain_shafts) return VX_E_UNSUPPORTED;
    bits = 0;
    if (spec_hi & 1u) bits |= RS_SKY_SPECULAR;
    if (!(spec_hi & 1048576u)) bits |= RS_LIGHTING_REPAIR;
    if (spec & 2147483648u) bits |= RS_SKY_TINT;
    if (spec_hi & 2048u) bits |= RS_SKY_COOL;
    if (spec_hi & 1024u) bits |= RS_SKY_LOOK;
    if (spec & 268435456u) bits |= RS_LIGHT_RGB;
    if (spec_hi & VX_FLAG_HI_EMITTER_WORLD) bits |= RS_EMITTER_GATHER;
    if (spec_hi & 8u) bits |= RS_SHORE_WET;
    if (spec & 16384u) bits |= RS_TEX_VARIATION;
    if (spec & VX_FLAG_GLASS) bits |= RS_GLASS;
    if (spec_hi & 32u) bits |= RS_WATER_SEC;
    if (spec_hi & 16u) bits |= RS_SHORE_FOAM;
    if (spec & 1048576u) bits |= RS_FLAT_SECONDARY;
    if (spec & VX_FLAG_DISTANT_SHADOWS) bits |= RS_DISTANT;
    if (spec & FLAG_TINT) bits |= RS_TINT;
    if (spec_hi & 512u) bits |= RS_TINT_BALANCE;
    if (spec & 67108864u) bits |= RS_PROBE_CUBE;
    if (spec & 134217728u) bits |= RS_PROBE_BOUNCE;
    if (spec & 536870912u) bits |= RS_PROBE_SUN;
    if (spec & 1073741824u) bits |= RS_PROBE_SUN_HIGH;
    if (spec_hi & 64u) bits |= RS_CAUSTICS;
    if (spec_hi & 65536u) bits |= RS_SOFT_SHADOWS;
    if (spec & 16777216u) bits |= RS_PROBE_TAP;
    if (spec & 33554432u) bits |= RS_PROBE_AMBIENT;
    return VX_OK;
}

}  // namespace

extern "C" {

int vx_resolve_set_assets(vx_ctx *c, const vx_resolve_assets *a) {
    if (!c || !a || !a->atlas || !a->block_faces || a->size == 0 || a->size > 4096 || a->layers == 0 ||
        a->layers > 65536 || a->mips == 0 || a->mips > 13 || (a->size >> (a->mips - 1u)) == 0 || a->n_faces == 0)
        return VX_E_ARG;
    size_t texels = 0;
    for (uint32_t l = 0; l < a->mips; l++) texels += (size_t)a->layers * (a->size >> l) * (a->size >> l);
    HIPX(hipSetDevice(c->device));
    int rc;
    if ((rc = vxh::ensure(c->atlas, texels * 4)) || (rc = vxh::ensure(c->faces, (size_t)a->n_faces * 4))) return rc;
    HIPX(hipMemcpyAsync(c->atlas.p, a->atlas, texels * 4, hipMemcpyHostToDevice, c->stream));
    HIPX(hipMemcpyAsync(c->faces.p, a->block_faces, (size_t)a->n_faces * 4, hipMemcpyHostToDevice, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    c->atlas_size = a->size;
    c->atlas_layers = a->layers;
    c->atlas_mips = a->mips;
    c->n_faces = a->n_faces;
    return VX_OK;
}

int vx_resolve(vx_ctx *c) {
    if (!c) return VX_E_ARG;
    if (!c->have_frame || !c->atlas_size) return VX_E_ARG;
    const vx_gpu_frame &g = c->gframe;
    uint32_t bits = 0;
    int rc = resolve_bits(c->spec_word, c->spec_hi_word, g, bits);
    if (rc) return rc;
    HIPX(hipSetDevice(c->device));
    if ((rc = vxh::ensure(c->color, (size_t)c->w * c->h * 16))) return rc;
    ResolveK k;
    k.p = vxh::make_pools(c);
    k.f = c->frame;
    k.fog_density = g.fog_density;
    k.fog_falloff = g.fog_falloff;
    k.ambient = g.ambient;
    k.fog_scatter = g.fog_scatter;
    k.fog_g = g.fog_g;
    k.fog_height = g.fog_height;
    k.mip_bias = g.mip_bias;
    k.haze_warm = g.haze_warm;
    k.zenith_deep = g.zenith_deep;
    k.bits = bits;
    k.atlas = (const uint32_t *)c->atlas.p;
    k.size = c->atlas_size;
    k.layers = c->atlas_layers;
    k.mips = c->atlas_mips;
    k.faces = (const uint32_t *)c->faces.p;
    k.n_faces = c->n_faces;
    k.vis = (const unsigned long long *)c->vis.p;
    k.sun = (const uint16_t *)c->shadow.p;
    k.heat = c->heat_valid ? (const uint32_t *)c->heat.p : nullptr;
    k.out = (float4 *)c->color.p;
    k.s = c->spec;
    k.water_absorb = g.water_absorb;
    k.hw = (c->w + 1u) / 2u;
    k.hh = (c->h + 1u) / 2u;
    k.shaft = nullptr;
    k.shaft_origin[0] = g.shaft_origin[0];
    k.shaft_origin[1] = g.shaft_origin[1];
    k.shaft_texel = g.shaft_texel;
    k.shaft_soft = g.shaft_soft;
    k.tint_strength = g.tint_strength;
    k.tint_freq_t = g.tint_freq_t;
    k.tint_freq_h = g.tint_freq_h;
    k.tint_seed = g.tint_seed;
    k.soft_h = nullptr;
    k.soft_mask = nullptr;
    if ((bits & RS_SOFT_SHADOWS) && c->shadow.p) {
        const size_t np = (size_t)c->w * c->h;
        if ((rc = vxh::ensure(c->soft_h, np * 4)) || (rc = vxh::ensure(c->soft_mask, np))) return rc;
        k.soft_h = (ushort2 *)c->soft_h.p;
        k.soft_mask = (uint8_t *)c->soft_mask.p;
    }
    k.probe = nullptr;
    if (bits & (RS_PROBE_CUBE | RS_PROBE_BOUNCE | RS_PROBE_TAP | RS_PROBE_AMBIENT)) {
        if ((rc = probe_buffer(c))) return rc;  // never filled: the field of ones
        k.probe = (const ushort4 *)c->probe.p;
    }
    if ((bits & RS_DISTANT) && (g.flags & VX_FLAG_DISTANT_SHADOWS)) {
        if ((rc = shaft_buffer(c))) return rc;  // never uploaded: the zeroed field
        k.shaft = (const float *)c->shaft.p + SHAFT_SLICE;
    }
    const bool lanes = !(g.flags & VX_FLAG_HEATMAP);
    if (lanes && (bits & RS_WATER_SEC)) {
        const size_t sec_bytes = (size_t)k.hw * k.hh * 8;
        if ((rc = vxh::ensure(c->sec_r, sec_bytes)) || (rc = vxh::ensure(c->sec_m, sec_bytes))) return rc;
        k.sec_r = (ushort4 *)c->sec_r.p;
        k.sec_m = (ushort4 *)c->sec_m.p;
        k_water_sec<<<((k.hw + 7u) / 8u) * ((k.hh + 7u) / 8u), 64, 0, c->stream>>>(k);
        VX_LAUNCHED();
    } else {
        k.sec_r = k.sec_m = nullptr;
    }
    if (k.soft_mask) {
        const unsigned blocks = (unsigned)(((size_t)c->w * c->h + 255) / 256);
        k_soft_h<<<blocks, 256, 0, c->stream>>>(k);
        k_soft_v<<<blocks, 256, 0, c->stream>>>(k);
        VX_LAUNCHED();
    }
    k_resolve<<<c->tx * c->ty, 64, 0, c->stream>>>(k);
    VX_LAUNCHED();
    if (lanes) {
        k_resolve_lanes<<<c->tx * c->ty, 64, 0, c->stream>>>(k);
        VX_LAUNCHED();
    }
    c->color_valid = true;
    c->image_valid = false;
    c->taa_valid = false;
    return VX_OK;
}

int vx_probe_fill(vx_ctx *c, const float *value, int noise) {
    if (!c) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    int rc;
    if ((rc = probe_buffer(c))) return rc;
    k_probe_fill<<<(unsigned)((PROBE_TEXELS + 255) / 256), 256, 0, c->stream>>>((ushort4 *)c->probe.p,
                                                                                value ? *value : 1.0f, noise ? 1 : 0);
    VX_LAUNCHED();
    c->probe_pinned = value != nullptr;
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_probe_write(vx_ctx *c, const int32_t origin[3], const float *occl, const float *bounce) {
    if (!c || !origin || !occl || !bounce) return VX_E_ARG;
    if ((origin[0] | origin[1] | origin[2]) & 63) return VX_E_ARG;  // a chunk origin
    HIPX(hipSetDevice(c->device));
    int rc;
    if ((rc = probe_buffer(c))) return rc;
    if (c->probe_pinned) return VX_OK;
    constexpr size_t N = 6u * 4096u;
    float *d = nullptr;
    HIPX(hipMalloc(&d, N * 4 * 4));
    HIPX(hipMemcpyAsync(d, occl, N * 4, hipMemcpyHostToDevice, c->stream));
    HIPX(hipMemcpyAsync(d + N, bounce, N * 12, hipMemcpyHostToDevice, c->stream));
    const int32_t sp = 4;
    auto rem = [](int32_t a, int32_t m) { return ((a % m) + m) % m; };
    const uint32_t bx = (uint32_t)rem(origin[0] / sp, (int32_t)PROBE_XZ);
    const uint32_t bz = (uint32_t)rem(origin[2] / sp, (int32_t)PROBE_XZ);
    const uint32_t by = (uint32_t)std::min(std::max(origin[1] / sp, 0), (int32_t)PROBE_Y - 16);
    k_probe_write<<<(unsigned)(N / 256), 256, 0, c->stream>>>((ushort4 *)c->probe.p, d, d + N, bx, by, bz);
    VX_LAUNCHED();
    HIPX(hipStreamSynchronize(c->stream));
    HIPX(hipFree(d));
    return VX_OK;
}

int vx_read_probe(vx_ctx *c, uint16_t *out, size_t n) {
    if (!c || !out || n != PROBE_TEXELS * 4) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    int rc;
    if ((rc = probe_buffer(c))) return rc;
    HIPX(hipMemcpyAsync(out, c->probe.p, PROBE_TEXELS * 8, hipMemcpyDeviceToHost, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_shaft_upload(vx_ctx *c, const float *heights, size_t n) {
    if (!c || !heights || n != SHAFT_SLICE) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    int rc;
    if ((rc = shaft_buffer(c))) return rc;
    HIPX(hipMemcpyAsync((float *)c->shaft.p + 2u * SHAFT_SLICE, heights, (size_t)SHAFT_SLICE * 4,
                        hipMemcpyHostToDevice, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_shaft_scan(vx_ctx *c, const vx_gpu_frame *frame) {
    if (!c || !frame) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    int rc;
    if ((rc = shaft_buffer(c))) return rc;
    const dim3 grid(SHAFT_DIM / 16u, SHAFT_DIM / 16u);
    for (uint32_t k = 0; k < SHAFT_STEPS; k++)
        k_shaft_scan<<<grid, 256, 0, c->stream>>>((float *)c->shaft.p, k, frame->sun_dir[0], frame->sun_dir[1],
                                                  frame->sun_dir[2], frame->shaft_texel);
    VX_LAUNCHED();
    return VX_OK;
}

int vx_read_shaft(vx_ctx *c, float *out, size_t n) {
    if (!c || !out || n != SHAFT_SLICE) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    int rc;
    if ((rc = shaft_buffer(c))) return rc;
    HIPX(hipMemcpyAsync(out, (const float *)c->shaft.p + SHAFT_SLICE, (size_t)SHAFT_SLICE * 4,
                        hipMemcpyDeviceToHost, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

// TAA history buffers: two f16 RGBA targets of the context's size, zeroed when (re)made.
static int taa_targets(vx_ctx *c) {
    const size_t bytes = (size_t)c->w * c->h * 8;
    if (c->hist[0].cap >= bytes && c->hist_len == bytes) return VX_OK;
    for (int i = 0; i < 2; i++) {
        int rc;
        if ((rc = vxh::ensure(c->hist[i], bytes))) return rc;
        HIPX(hipMemsetAsync(c->hist[i].p, 0, bytes, c->stream));
    }
    c->hist_len = bytes;
    c->taa_last = 0;
    return VX_OK;
}

int vx_taa(vx_ctx *c) {
    if (!c || !c->color_valid || !c->have_frame) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    int rc;
    if ((rc = taa_targets(c))) return rc;
    const vx_gpu_frame &g = c->gframe;
    TaaK k;
    k.p = vxh::make_pools(c);
    k.f = c->frame;
    std::memcpy(k.prev, g.prev_view_proj, sizeof k.prev);
    k.feedback = g.taa_feedback;
    k.vis = (const unsigned long long *)c->vis.p;
    k.src = (const float4 *)c->color.p;
    k.hist = (const ushort4 *)c->hist[c->taa_last].p;
    k.out = (ushort4 *)c->hist[c->taa_last ^ 1u].p;
    k_taa<<<c->tx * c->ty, 64, 0, c->stream>>>(k);
    VX_LAUNCHED();
    c->taa_last ^= 1u;
    c->taa_valid = true;
    c->image_valid = false;
    return VX_OK;
}

int vx_read_taa(vx_ctx *c, uint16_t *rgba, size_t n) {
    if (!c || !rgba || n != (size_t)c->w * c->h * 4 || !c->taa_valid || c->hist_len != n * 2) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    HIPX(hipMemcpyAsync(rgba, c->hist[c->taa_last].p, n * 2, hipMemcpyDeviceToHost, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_write_taa_history(vx_ctx *c, const uint16_t *rgba, size_t n) {
    if (!c || !rgba || n != (size_t)c->w * c->h * 4) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    int rc;
    if ((rc = taa_targets(c))) return rc;
    HIPX(hipMemcpyAsync(c->hist[c->taa_last].p, rgba, n * 2, hipMemcpyHostToDevice, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_present(vx_ctx *c, const vx_present_desc *d) {
    if (!c || !d || !c->color_valid || d->tone > 1u || d->grade > 2u || d->encode > 1u) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    const uint32_t n = c->w * c->h;
    int rc;
    if ((rc = vxh::ensure(c->image, (size_t)n * 4))) return rc;
    const float4 *src = (const float4 *)c->color.p;
    if (c->taa_valid) {
        // After vx_taa the blit draws the TAA output (src/render/mod.rs:1908, last_blit).
        if ((rc = vxh::ensure(c->taa_f32, (size_t)n * 16))) return rc;
        k_half_to_float<<<(n + 255) / 256, 256, 0, c->stream>>>((const ushort4 *)c->hist[c->taa_last].p, n,
                                                                (float4 *)c->taa_f32.p);
        VX_LAUNCHED();
        src = (const float4 *)c->taa_f32.p;
    }
    k_present<<<(n + 255) / 256, 256, 0, c->stream>>>(src, n, d->encode, d->tone, d->grade,
                                                      d->strength, (uint32_t *)c->image.p);
    VX_LAUNCHED();
    c->image_valid = true;
    return VX_OK;
}

int vx_read_image(vx_ctx *c, uint8_t *rgba, size_t n) {
    if (!c || !rgba || n != (size_t)c->w * c->h * 4 || !c->image_valid) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    HIPX(hipMemcpyAsync(rgba, c->image.p, n, hipMemcpyDeviceToHost, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_read_color(vx_ctx *c, float *rgba, size_t n) {
    if (!c || !rgba || n != (size_t)c->w * c->h * 4 || !c->color_valid) return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    HIPX(hipMemcpyAsync(rgba, c->color.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

}  // extern "C"
