// probe.cpp -- the ambient probe bake (src/probe.rs:1-379).
//
// Per LOD-0 chunk, a 16^3 lattice of probes 4 blocks apart. Each probe column looks along
// 16 azimuths at 20 radii (1 to 192 blocks, geometric): the terrain heights there give
// the horizon tangent seen from every probe height, and the surface blocks' albedo
// (every second radius, weighted by distance and, with bounce_shadow, by how exposed
// that ground is) gives the colour the ground bounces. Per probe and cube axis the bake
// stores the sky occlusion (a cosine lobe over the horizon elevations, faded to 1 where
// the probe is buried) and the bounce colour (the lobe-weighted albedo, normalised so
// that grass ground leaves the floor light's luminance alone).
#include <algorithm>
#include <cmath>

#include "vxhost.h"

namespace {

constexpr int32_t SPACING = 4, PER_AXIS = 16, AXES = 6, PROBES = PER_AXIS * PER_AXIS * PER_AXIS;
constexpr int32_t STEPS = 20, AZIMUTHS = 16, ALBEDO_STRIDE = 2;
constexpr float REACH = 192.0f, SKY_BLEND = 0.25f, UPLOAD_REACH = 224.0f;
constexpr float FLOOR_TINT[3] = {0.85f, 0.90f, 1.00f};
constexpr uint16_t GRASS = 3, BLOCK_COUNT = 26;
// block::ALBEDO (src/block.rs:187-214): the linear mean colour of each block's texture
constexpr float ALBEDO[BLOCK_COUNT][3] = {
    {0.1738f, 0.1738f, 0.1738f}, {0.1738f, 0.1738f, 0.1738f}, {0.2066f, 0.1016f, 0.0470f},
    {0.0920f, 0.2947f, 0.0407f}, {0.6170f, 0.5245f, 0.2573f}, {0.1598f, 0.1458f, 0.1328f},
    {0.2828f, 0.1724f, 0.0622f}, {0.0389f, 0.1941f, 0.0226f}, {0.3439f, 0.1974f, 0.0648f},
    {0.2269f, 0.2269f, 0.2269f}, {0.4397f, 0.6282f, 0.7365f}, {0.8058f, 0.6399f, 0.2432f},
    {0.0651f, 0.0651f, 0.0744f}, {0.0097f, 0.0931f, 0.1470f}, {0.7240f, 0.7812f, 0.9029f},
    {0.0844f, 0.0386f, 0.0137f}, {0.0153f, 0.0848f, 0.0386f}, {0.1574f, 0.1857f, 0.1178f},
    {0.0121f, 0.0685f, 0.0063f}, {0.0317f, 0.0936f, 0.0152f}, {0.6480f, 0.2539f, 0.0576f},
    {0.1397f, 0.2884f, 0.6801f}, {0.2356f, 0.6124f, 0.1454f}, {0.0317f, 0.0936f, 0.0152f},
    {0.0371f, 0.0987f, 0.0080f}, {0.0266f, 0.0444f, 0.0064f}};

float luma(const float c[3]) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }

struct Tables {
    float dir[AZIMUTHS][2];
    float lobe[AXES][AZIMUTHS];
    float lobe_sum[AXES];
};

// probe::tables: the azimuth directions and each horizontal axis' cosine lobe over them
Tables tables() {
    Tables t{};
    for (int k = 0; k < AZIMUTHS; k++) {
        const float a = 6.28318530717958647692f * (float)k / (float)AZIMUTHS;
        t.dir[k][0] = std::cos(a);
        t.dir[k][1] = std::sin(a);
    }
    for (int axis = 0; axis < AXES; axis++) {
        float ax = 0.0f, az = 0.0f;
        if (axis == 0) ax = -1.0f;
        else if (axis == 1) ax = 1.0f;
        else if (axis == 4) az = -1.0f;
        else if (axis == 5) az = 1.0f;
        for (int k = 0; k < AZIMUTHS; k++) {
            const float w = std::max(t.dir[k][0] * ax + t.dir[k][1] * az, 0.0f);
            t.lobe[axis][k] = w;
            t.lobe_sum[axis] += w;
        }
    }
    return t;
}

// cube_from_horizon: per axis, the sky the horizon leaves (elevation-weighted lobe for the
// horizontal axes, the open fraction for +Y, 1 for -Y), faded towards 1 as the probe's
// open fraction drops below SKY_BLEND; returns the fade ("believe") too.
float cube_from_horizon(const float tan_h[AZIMUTHS], const Tables &tb, float out[AXES]) {
    constexpr float QUARTER_PI = 0.785398163397448309616f;
    float up = 0.0f, elev[AZIMUTHS];
    for (int k = 0; k < AZIMUTHS; k++) {
        const float t = std::max(tan_h[k], 0.0f);
        up += 1.0f / (1.0f + t * t);
        elev[k] = QUARTER_PI - 0.5f * std::atan(t) - 0.5f * t / (1.0f + t * t);
    }
    up /= (float)AZIMUTHS;
    const float tt = std::min(std::max(up / SKY_BLEND, 0.0f), 1.0f);
    const float believe = tt * tt * (3.0f - 2.0f * tt);
    for (int axis = 0; axis < AXES; axis++) {
        float occl;
        if (axis == 3) {
            occl = up;
        } else if (axis == 2) {
            occl = 1.0f;
        } else {
            float num = 0.0f;
            for (int k = 0; k < AZIMUTHS; k++) num += tb.lobe[axis][k] * elev[k];
            occl = num / (tb.lobe_sum[axis] * QUARTER_PI);
        }
        out[axis] = 1.0f + (occl - 1.0f) * believe;
    }
    return believe;
}

// bounce_lobes: the vertical axes gather every azimuth
void bounce_lobes(const Tables &tb, float out[AXES][AZIMUTHS]) {
    for (int a = 0; a < AXES; a++)
        for (int k = 0; k < AZIMUTHS; k++) out[a][k] = (a == 2 || a == 3) ? 1.0f : tb.lobe[a][k];
}

// bounce_scale: grass ground (BOUNCE_REF) maps to the floor tint's luminance
void bounce_scale(float s[3]) {
    const float gain = luma(ALBEDO[GRASS]);
    for (int c = 0; c < 3; c++) s[c] = luma(FLOOR_TINT) / (gain * FLOOR_TINT[c]);
}

void bounce_from_albedo(const float alb[AZIMUTHS][3], const float alb_w[AZIMUTHS], const float blobe[AXES][AZIMUTHS],
                        const float scale[3], float believe, float out[AXES][3]) {
    for (int axis = 0; axis < AXES; axis++) {
        out[axis][0] = out[axis][1] = out[axis][2] = 1.0f;
        float num[3] = {0.0f, 0.0f, 0.0f}, den = 0.0f;
        for (int k = 0; k < AZIMUTHS; k++) {
            const float w = blobe[axis][k];
            if (w <= 0.0f) continue;
            for (int c = 0; c < 3; c++) num[c] += w * alb[k][c];
            den += w * alb_w[k];
        }
        if (den <= 0.0f) continue;
        const float inv = 1.0f / den;
        for (int c = 0; c < 3; c++) out[axis][c] = 1.0f + (num[c] * inv * scale[c] - 1.0f) * believe;
    }
}

// ray_exposure: how open each gathered ground sample is to the sky along the ray, from the
// steepest rise beyond it and before it
void ray_exposure(const float col_h[STEPS], const float radius[STEPS], float out[STEPS]) {
    for (int si = 0; si < STEPS; si++) {
        out[si] = 1.0f;
        if (si % ALBEDO_STRIDE != 0) continue;
        const float h = col_h[si];
        float t_out = 0.0f, t_in = 0.0f;
        for (int sj = 0; sj < STEPS; sj++) {
            if (sj == si) continue;
            const float d = std::max(std::fabs(radius[sj] - radius[si]), 1.0f);
            const float t = (col_h[sj] - h) / d;
            if (sj > si) t_out = std::max(t_out, t);
            else t_in = std::max(t_in, t);
        }
        out[si] = 0.5f * (1.0f / (1.0f + t_out * t_out) + 1.0f / (1.0f + t_in * t_in));
    }
}

}  // namespace

extern "C" {

// probe::bake with this generator's heights (vxw_height) and surface blocks
int vxw_probe_bake(const vxw_gen_params *gen, const int32_t origin[3], int bounce_shadow, float *occl,
                   float *bounce) {
    if (!gen || !origin || !occl || !bounce) return VXW_E_ARG;
    std::fill(occl, occl + AXES * PROBES, 0.0f);
    std::fill(bounce, bounce + AXES * PROBES * 3, 1.0f);
    const Tables tb = tables();
    float blobe[AXES][AZIMUTHS], scale[3];
    bounce_lobes(tb, blobe);
    bounce_scale(scale);
    float radius[STEPS];
    const float ratio = std::pow(REACH / 1.0f, 1.0f / (float)(STEPS - 1));
    float r = 1.0f;
    for (int i = 0; i < STEPS; i++) {
        radius[i] = r;
        r *= ratio;
    }
    static thread_local float tan_h[PER_AXIS][AZIMUTHS], alb[PER_AXIS][AZIMUTHS][3], alb_w[PER_AXIS][AZIMUTHS];
    float col_h[STEPS], col_a[STEPS][3], lit[STEPS];
    for (int pz = 0; pz < PER_AXIS; pz++)
        for (int px = 0; px < PER_AXIS; px++) {
            const int32_t cx = origin[0] + SPACING / 2 + SPACING * px;
            const int32_t cz = origin[2] + SPACING / 2 + SPACING * pz;
            for (int py = 0; py < PER_AXIS; py++)
                for (int k = 0; k < AZIMUTHS; k++) {
                    tan_h[py][k] = 0.0f;
                    alb[py][k][0] = alb[py][k][1] = alb[py][k][2] = 0.0f;
                    alb_w[py][k] = 0.0f;
                }
            for (int k = 0; k < AZIMUTHS; k++) {
                const float dx = tb.dir[k][0], dz = tb.dir[k][1];
                for (int si = 0; si < STEPS; si++) {
                    const int32_t wx = cx + (int32_t)std::round(dx * radius[si]);
                    const int32_t wz = cz + (int32_t)std::round(dz * radius[si]);
                    const int32_t hi = vxw_height(gen, wx, wz);
                    col_h[si] = (float)hi;
                    if (si % ALBEDO_STRIDE == 0) {
                        const uint16_t b = std::min<uint16_t>(vxw_surface_block(gen, wx, wz, hi), BLOCK_COUNT - 1);
                        for (int c = 0; c < 3; c++) col_a[si][c] = ALBEDO[b][c];
                    } else {
                        col_a[si][0] = col_a[si][1] = col_a[si][2] = 0.0f;
                    }
                }
                if (bounce_shadow) ray_exposure(col_h, radius, lit);
                else std::fill(lit, lit + STEPS, 1.0f);
                for (int si = 0; si < STEPS; si++) {
                    const float h = col_h[si], rr = radius[si];
                    const bool gathering = si % ALBEDO_STRIDE == 0;
                    const float a[3] = {col_a[si][0] * lit[si], col_a[si][1] * lit[si], col_a[si][2] * lit[si]};
                    for (int py = 0; py < PER_AXIS; py++) {
                        const float y = (float)(origin[1] + SPACING / 2 + SPACING * py);
                        const float t = (h - y) / rr;
                        if (t > tan_h[py][k]) tan_h[py][k] = t;
                        if (!gathering) continue;
                        const float d = std::max(std::fabs(y - h), 1.0f);
                        const float dr = d * rr;
                        const float s = d * d + rr * rr;
                        const float w = (dr * dr) / (s * s);
                        for (int c = 0; c < 3; c++) alb[py][k][c] += w * a[c];
                        alb_w[py][k] += w;
                    }
                }
            }
            for (int py = 0; py < PER_AXIS; py++) {
                float cube[AXES], b[AXES][3];
                const float believe = cube_from_horizon(tan_h[py], tb, cube);
                bounce_from_albedo(alb[py], alb_w[py], blobe, scale, believe, b);
                const int i = (pz * PER_AXIS + py) * PER_AXIS + px;
                for (int axis = 0; axis < AXES; axis++) {
                    occl[axis * PROBES + i] = cube[axis];
                    const int o = (axis * PROBES + i) * 3;
                    bounce[o] = b[axis][0];
                    bounce[o + 1] = b[axis][1];
                    bounce[o + 2] = b[axis][2];
                }
            }
        }
    return VXW_OK;
}

void vxw_probe_cube_from_tangents(const float tan_h[16], float out[6]) {
    if (!tan_h || !out) return;
    const Tables tb = tables();
    (void)cube_from_horizon(tan_h, tb, out);
}

void vxw_probe_bounce_of_uniform_ground(const float albedo[3], float out[18]) {
    if (!albedo || !out) return;
    const Tables tb = tables();
    float blobe[AXES][AZIMUTHS], scale[3], alb[AZIMUTHS][3], w[AZIMUTHS], b[AXES][3];
    bounce_lobes(tb, blobe);
    bounce_scale(scale);
    for (int k = 0; k < AZIMUTHS; k++) {
        alb[k][0] = albedo[0];
        alb[k][1] = albedo[1];
        alb[k][2] = albedo[2];
        w[k] = 1.0f;
    }
    bounce_from_albedo(alb, w, blobe, scale, 1.0f, b);
    for (int a = 0; a < AXES; a++)
        for (int c = 0; c < 3; c++) out[a * 3 + c] = b[a][c];
}

int vxw_probe_uploadable(const int32_t origin[3], const float cam[3]) {
    if (!origin || !cam) return 0;
    const float dx = (float)origin[0] - cam[0], dz = (float)origin[2] - cam[2];
    return dx * dx + dz * dz <= UPLOAD_REACH * UPLOAD_REACH ? 1 : 0;
}

float vxw_albedo(uint16_t id, int c) { return id < BLOCK_COUNT && c >= 0 && c < 3 ? ALBEDO[id][c] : 0.0f; }

}  // extern "C"
