// shaft.cpp -- the height window of the distant terrain shadows (src/shaft.rs:1-154).
//
// A 512 x 512 window of column heights, one texel per `texel` x `texel` blocks, centred on
// the camera's column. The GPU scans it towards the sun (vx_shaft_scan) and the resolve
// darkens the sun term of points whose scanned horizon lies above them
// (terrain_shade_beyond). When the camera moves to another texel the window scrolls: the
// columns still inside keep their heights and the new ones are sampled from the
// generator at texel centres; an unmoved camera refills nothing.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "vxhost.h"

struct vxw_shaft {
    std::vector<float> heights;
    int32_t origin[2] = {0, 0};  // texel coordinates of the window's first column
    int32_t texel = VXW_SHAFT_TEXEL;
    bool filled = false, dirty = false;
};

namespace {

constexpr int32_t DIM = (int32_t)VXW_SHAFT_DIM;

// ShaftField::sample (src/shaft.rs:134-153): bilinear at a fractional texel, clamped.
float sample(const vxw_shaft *s, float x, float z) {
    const float hi = (float)(DIM - 1);
    const float cx = std::fmin(std::fmax(x, 0.0f), hi), cz = std::fmin(std::fmax(z, 0.0f), hi);
    const float x0f = std::floor(cx), z0f = std::floor(cz);
    const float fx = cx - x0f, fz = cz - z0f;
    const int32_t x0 = (int32_t)x0f, z0 = (int32_t)z0f;
    const int32_t x1 = std::min(x0 + 1, DIM - 1), z1 = std::min(z0 + 1, DIM - 1);
    const float *h = s->heights.data();
    const float a = h[z0 * DIM + x0], b = h[z0 * DIM + x1], c = h[z1 * DIM + x0], d = h[z1 * DIM + x1];
    const float top = a + (b - a) * fx;
    const float bot = c + (d - c) * fx;
    return top + (bot - top) * fz;
}

}  // namespace

extern "C" {

vxw_shaft *vxw_shaft_new(int32_t texel) {
    if (texel < 1 || texel > 64) return nullptr;  // ShaftField::with_texel: past the scan's reach
    vxw_shaft *s = new (std::nothrow) vxw_shaft();
    if (!s) return nullptr;
    s->heights.assign((size_t)DIM * DIM, 0.0f);
    s->texel = texel;
    return s;
}

void vxw_shaft_free(vxw_shaft *s) { delete s; }

// ShaftField::update (src/shaft.rs:58-107), the generator's column height for
// WorldGen::height; canopy_lift (tree-line canopies) is not restated.
int vxw_shaft_update(vxw_shaft *s, const vxw_gen_params *gen, float cam_x, float cam_z, float canopy_lift) {
    if (!s || !gen || canopy_lift > 0.0f) return VXW_E_ARG;
    const int32_t half = DIM / 2;
    const int32_t want[2] = {(int32_t)std::floor(cam_x / (float)s->texel) - half,
                             (int32_t)std::floor(cam_z / (float)s->texel) - half};
    s->dirty = false;
    if (s->filled && want[0] == s->origin[0] && want[1] == s->origin[1]) return VXW_OK;
    const int32_t d[2] = {want[0] - s->origin[0], want[1] - s->origin[1]};
    std::vector<float> next((size_t)DIM * DIM);
    for (int32_t zz = 0; zz < DIM; zz++) {
        const int32_t oz = zz + d[1];
        for (int32_t xx = 0; xx < DIM; xx++) {
            const int32_t ox = xx + d[0];
            const bool carried = s->filled && oz >= 0 && oz < DIM && ox >= 0 && ox < DIM;
            if (carried) {
                next[(size_t)zz * DIM + xx] = s->heights[(size_t)oz * DIM + ox];
            } else {
                const int32_t wx = (want[0] + xx) * s->texel + s->texel / 2;
                const int32_t wz = (want[1] + zz) * s->texel + s->texel / 2;
                next[(size_t)zz * DIM + xx] = (float)vxw_height(gen, wx, wz);
            }
        }
    }
    s->heights.swap(next);
    s->origin[0] = want[0];
    s->origin[1] = want[1];
    s->filled = true;
    s->dirty = true;
    return VXW_OK;
}

const float *vxw_shaft_heights(const vxw_shaft *s) { return s ? s->heights.data() : nullptr; }

void vxw_shaft_origin_world(const vxw_shaft *s, float out[2]) {
    if (!s || !out) return;
    out[0] = (float)(s->origin[0] * s->texel);
    out[1] = (float)(s->origin[1] * s->texel);
}

int32_t vxw_shaft_texel(const vxw_shaft *s) { return s ? s->texel : 0; }

int vxw_shaft_dirty(const vxw_shaft *s) { return s && s->dirty ? 1 : 0; }

// ShaftField::envelope_reference (src/shaft.rs:117-132): the highest sampled height along
// the sun's horizontal direction over 2^STEPS texels, each lowered by the sun's climb.
float vxw_shaft_envelope(const vxw_shaft *s, const float sun[3], uint32_t tx, uint32_t tz) {
    if (!s || !sun || tx >= (uint32_t)DIM || tz >= (uint32_t)DIM) return NAN;
    const float horiz = std::fmax(std::sqrt(sun[0] * sun[0] + sun[2] * sun[2]), 1e-3f);
    const float dir[2] = {sun[0] / horiz, sun[2] / horiz};
    const float slope = sun[1] / horiz;
    float best = s->heights[(size_t)tz * DIM + tx];
    const uint32_t reach = 1u << VXW_SHAFT_STEPS;
    for (uint32_t m = 1; m <= reach; m++) {
        const float sx = (float)tx + dir[0] * (float)m;
        const float sz = (float)tz + dir[1] * (float)m;
        const float h = sample(s, sx, sz);
        best = std::fmax(best, h - (float)m * (float)s->texel * slope);
    }
    return best;
}

}  // extern "C"
