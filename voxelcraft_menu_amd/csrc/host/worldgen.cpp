// worldgen.cpp -- seeded terrain for the benchmark worlds (fixture support).
//
// Structure follows the reference generator (src/worldgen.rs:139-533): a column
// height from three fBm layers combined as base + hills + mountains + detail
// (src/worldgen.rs:199-211), column materials (bedrock / stone / dirt / surface,
// sand on beaches, water up to the sea level, src/worldgen.rs:214-256), 3D-noise
// caves, and seeded trees. The noise itself is this project's own gradient noise,
// not FastNoiseLite, so the grids are not bit-identical to the reference worlds;
// the march is compared against the oracle on the same grid, which is what parity
// is defined on.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "vxhost.h"
#include "world.hpp"

namespace vxh {

static inline uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

static inline uint32_t hash3i(int32_t x, int32_t y, int32_t z, int32_t seed) {
    uint32_t h = (uint32_t)x * 0x8da6b343u ^ (uint32_t)y * 0xd8163841u ^ (uint32_t)z * 0xcb1ab31fu ^
                 (uint32_t)seed * 0x165667b1u;
    return mix32(h);
}

static inline float unit01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

static inline float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }

// 2D gradient noise: 8 unit gradients picked by a lattice hash, quintic fade.
static float grad2(int32_t ix, int32_t iz, int32_t seed, float dx, float dz) {
    static const float G[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1},
                                  {0.70710678f, 0.70710678f}, {-0.70710678f, 0.70710678f},
                                  {0.70710678f, -0.70710678f}, {-0.70710678f, -0.70710678f}};
    uint32_t h = hash3i(ix, 0x5bd1, iz, seed) & 7u;
    return G[h][0] * dx + G[h][1] * dz;
}

static float noise2(int32_t seed, float x, float z) {
    float fx = std::floor(x), fz = std::floor(z);
    int32_t ix = (int32_t)fx, iz = (int32_t)fz;
    float dx = x - fx, dz = z - fz;
    float u = fade(dx), v = fade(dz);
    float a = grad2(ix, iz, seed, dx, dz), b = grad2(ix + 1, iz, seed, dx - 1, dz);
    float c = grad2(ix, iz + 1, seed, dx, dz - 1), d = grad2(ix + 1, iz + 1, seed, dx - 1, dz - 1);
    float r = (a + (b - a) * u) + ((c + (d - c) * u) - (a + (b - a) * u)) * v;
    return r * 1.41421356f;
}

static float grad3(int32_t ix, int32_t iy, int32_t iz, int32_t seed, float dx, float dy, float dz) {
    static const float G[12][3] = {{1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1},
                                   {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
    uint32_t h = hash3i(ix, iy, iz, seed) % 12u;
    return G[h][0] * dx + G[h][1] * dy + G[h][2] * dz;
}

static float noise3(int32_t seed, float x, float y, float z) {
    float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    float dx = x - fx, dy = y - fy, dz = z - fz;
    float u = fade(dx), v = fade(dy), w = fade(dz);
    float c[2][2][2];
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++)
                c[k][j][i] = grad3(ix + i, iy + j, iz + k, seed, dx - i, dy - j, dz - k);
    float x00 = c[0][0][0] + (c[0][0][1] - c[0][0][0]) * u, x10 = c[0][1][0] + (c[0][1][1] - c[0][1][0]) * u;
    float x01 = c[1][0][0] + (c[1][0][1] - c[1][0][0]) * u, x11 = c[1][1][0] + (c[1][1][1] - c[1][1][0]) * u;
    float y0 = x00 + (x10 - x00) * v, y1 = x01 + (x11 - x01) * v;
    return (y0 + (y1 - y0) * w) * 1.1f;
}

static float fbm2(int32_t seed, float freq, int octaves, float x, float z) {
    float sum = 0, amp = 1, norm = 0;
    for (int o = 0; o < octaves; o++) {
        sum += noise2(seed + o * 7919, x * freq, z * freq) * amp;
        norm += amp;
        amp *= 0.5f;
        freq *= 2.0f;
    }
    return sum / norm;
}

struct Gen {
    vxw_gen_params p;
    float cave_q = 1e9f;  // carve where cave noise > cave_q

    explicit Gen(const vxw_gen_params &pp) : p(pp) {
        if (p.cave_fraction > 0.0f) {
            // Threshold at the (1 - fraction) quantile of the cave noise, sampled once.
            std::vector<float> s;
            s.reserve(20000);
            for (int i = 0; i < 20000; i++) {
                uint32_t h = mix32((uint32_t)i * 2654435761u + (uint32_t)p.seed);
                float x = unit01(h) * 4096.f, y = unit01(mix32(h ^ 0x1234567u)) * 512.f,
                      z = unit01(mix32(h ^ 0x89abcdefu)) * 4096.f;
                s.push_back(cave(x, y, z));
            }
            std::sort(s.begin(), s.end());
            float f = std::min(std::max(p.cave_fraction, 0.0f), 0.95f);
            cave_q = s[(size_t)((1.0f - f) * (float)(s.size() - 1))];
        }
    }

    float cave(float x, float y, float z) const {
        return noise3(p.seed + 4, x * 0.028f, y * 0.042f, z * 0.028f) * 0.7f +
               noise3(p.seed + 5, x * 0.061f, y * 0.061f, z * 0.061f) * 0.3f;
    }

    int32_t height(int32_t x, int32_t z) const {
        float fx = (float)x, fz = (float)z;
        float c = fbm2(p.seed, 0.0008f, 3, fx, fz);
        float h = fbm2(p.seed + 1, 0.0022f, 3, fx, fz);
        float d = fbm2(p.seed + 2, 0.020f, 2, fx, fz);
        // Terrain sits around 8 above the sea with lakes in the lows.
        float base = (float)p.sea_level + 8.0f + c * 40.0f;
        float mask = std::min(std::max((c - 0.05f) * 2.4f, 0.0f), 1.0f);
        float peaks = std::max(h, 0.0f);
        float mountains = peaks * peaks * 90.0f * mask;
        float ht = base + h * 24.0f + mountains + d * 2.2f;
        return std::min(std::max((int32_t)std::lround(ht), 2), WORLD_HEIGHT - 8);
    }

    BlockId column_block(int32_t y, int32_t h) const {
        if (y >= h) return y < p.sea_level ? WATER : AIR;
        if (y == 0) return BEDROCK;
        bool beach = h <= p.sea_level + 2;
        bool snowy = h - 1 >= 196;
        if (y == h - 1) return beach ? SAND : (snowy ? SNOW : GRASS);
        if (y >= h - 4) return beach ? SAND : DIRT;
        return STONE;
    }

    void chunk(int32_t cx, int32_t cy, int32_t cz, BlockId *out) const {
        std::fill(out, out + VOL, AIR);
        int32_t ox = cx * 64, oy = cy * 64, oz = cz * 64;
        static thread_local std::vector<int32_t> hs;
        hs.assign(64 * 64, 0);
        for (int z = 0; z < 64; z++)
            for (int x = 0; x < 64; x++) hs[x + z * 64] = height(ox + x, oz + z);
        // Cave field: noise on a 4-voxel lattice, trilinear in between.
        static thread_local std::vector<float> lat;
        bool caves = p.cave_fraction > 0.0f;
        if (caves) {
            lat.assign(17 * 17 * 17, 0.f);
            for (int k = 0; k < 17; k++)
                for (int j = 0; j < 17; j++)
                    for (int i = 0; i < 17; i++)
                        lat[i + 17 * (j + 17 * k)] = cave((float)(ox + 4 * i), (float)(oy + 4 * j), (float)(oz + 4 * k));
        }
        for (int z = 0; z < 64; z++)
            for (int x = 0; x < 64; x++) {
                int32_t h = hs[x + z * 64];
                int32_t wx = ox + x, wz = oz + z;
                uint32_t ch = hash3i(wx, 77, wz, p.seed);
                bool mouth = unit01(hash3i(wx >> 3, 91, wz >> 3, p.seed)) < 0.08f;
                int32_t crust = mouth ? 0 : 4;
                for (int y = 0; y < 64; y++) {
                    int32_t wy = oy + y;
                    BlockId b = column_block(wy, h);
                    if (caves && b != AIR && b != WATER && b != BEDROCK && wy < h - crust && wy > 0) {
                        float fx = x * 0.25f, fy = y * 0.25f, fz = z * 0.25f;
                        int i = x >> 2, j = y >> 2, k = z >> 2;
                        float u = fx - i, v = fy - j, w = fz - k;
                        auto L = [&](int a, int bb, int c) { return lat[(i + a) + 17 * ((j + bb) + 17 * (k + c))]; };
                        float n = ((L(0, 0, 0) * (1 - u) + L(1, 0, 0) * u) * (1 - v) +
                                   (L(0, 1, 0) * (1 - u) + L(1, 1, 0) * u) * v) * (1 - w) +
                                  ((L(0, 0, 1) * (1 - u) + L(1, 0, 1) * u) * (1 - v) +
                                   (L(0, 1, 1) * (1 - u) + L(1, 1, 1) * u) * v) * w;
                        if (n > cave_q) b = AIR;
                    }
                    if (wy == h - 1 && b != AIR && b != WATER && h > p.sea_level) {
                        float r = unit01(ch);
                        if (r < p.glass_fraction) b = GLASS;
                        else if (r < p.glass_fraction + p.leaf_extra) b = LEAVES;
                    }
                    out[dense_index((uint32_t)x, (uint32_t)y, (uint32_t)z)] = b;
                }
            }
        // Trees: anchors in columns near the chunk; trunk 4..6 logs, leaf ball radius 2.
        if (p.tree_density > 0.0f) {
            for (int32_t az = oz - 3; az < oz + 67; az++)
                for (int32_t ax = ox - 3; ax < ox + 67; ax++) {
                    uint32_t th = hash3i(ax, 13, az, p.seed);
                    if (unit01(th) >= p.tree_density) continue;
                    int32_t h = height(ax, az);
                    if (h <= p.sea_level + 1 || h - 1 >= 196) continue;
                    int32_t trunk = 4 + (int32_t)((th >> 3) % 3u);
                    int32_t top = h + trunk;
                    bool pine = ((th >> 7) & 3u) == 0u;
                    for (int32_t y = top - 2; y <= top + 2; y++)
                        for (int32_t z = az - 2; z <= az + 2; z++)
                            for (int32_t x = ax - 2; x <= ax + 2; x++) {
                                int32_t dx = x - ax, dy = y - top, dz = z - az;
                                if (dx * dx + dy * dy + dz * dz > 5) continue;
                                int32_t lx = x - ox, ly = y - oy, lz = z - oz;
                                if (lx < 0 || ly < 0 || lz < 0 || lx >= 64 || ly >= 64 || lz >= 64) continue;
                                BlockId &cell = out[dense_index((uint32_t)lx, (uint32_t)ly, (uint32_t)lz)];
                                if (cell == AIR) cell = pine ? PINE_LEAVES : LEAVES;
                            }
                    for (int32_t y = h; y < top; y++) {
                        int32_t lx = ax - ox, ly = y - oy, lz = az - oz;
                        if (lx < 0 || ly < 0 || lz < 0 || lx >= 64 || ly >= 64 || lz >= 64) continue;
                        out[dense_index((uint32_t)lx, (uint32_t)ly, (uint32_t)lz)] = LOG;
                    }
                }
        }
    }

    // A coarse stand-in (lod > 0) with generate_coarse's structure (src/worldgen.rs:
    // 535-562): one height sample per coarse column, each cell given the block at the
    // top of its span below the column's ceiling (the terrain or the sea surface). The
    // surface cell takes the LOD-0 column's glass / leaf pick; caves and trees are
    // left out of stand-ins.
    void coarse(const Key &k, BlockId *out) const {
        std::fill(out, out + VOL, AIR);
        const int32_t s = k.voxel_size(), span = k.size();
        const int32_t ox = k.x * span, oy = k.y * span, oz = k.z * span;
        for (int z = 0; z < 64; z++)
            for (int x = 0; x < 64; x++) {
                const int32_t wx = ox + x * s, wz = oz + z * s;
                const int32_t h = height(wx, wz);
                const int32_t ceiling = std::max(h, p.sea_level);
                const float r = unit01(hash3i(wx, 77, wz, p.seed));
                for (int y = 0; y < 64; y++) {
                    const int32_t wy = oy + y * s;
                    if (wy >= ceiling) break;
                    const int32_t top = std::min(wy + s - 1, ceiling - 1);
                    BlockId b = column_block(top, h);
                    if (top == h - 1 && b != AIR && b != WATER && h > p.sea_level) {
                        if (r < p.glass_fraction) b = GLASS;
                        else if (r < p.glass_fraction + p.leaf_extra) b = LEAVES;
                    }
                    out[dense_index((uint32_t)x, (uint32_t)y, (uint32_t)z)] = b;
                }
            }
    }
};

}  // namespace vxh

using namespace vxh;

extern "C" int32_t vxw_height(const vxw_gen_params *p, int32_t x, int32_t z) {
    if (!p) return 0;
    vxw_gen_params q = *p;
    q.cave_fraction = 0.0f;
    Gen g(q);
    return g.height(x, z);
}

extern "C" uint16_t vxw_surface_block(const vxw_gen_params *p, int32_t x, int32_t z, int32_t h) {
    if (!p) return 0;
    vxw_gen_params q = *p;
    q.cave_fraction = 0.0f;
    Gen g(q);
    (void)x;
    (void)z;
    return g.column_block(std::max(h, q.sea_level) - 1, h);  // the column top: the terrain or the sea surface
}

extern "C" int vxw_generate_dense(const vxw_gen_params *p, int32_t cx, int32_t cy, int32_t cz, uint16_t *out) {
    if (!p || !out) return VXW_E_ARG;
    Gen g(*p);
    g.chunk(cx, cy, cz, out);
    return VXW_OK;
}

extern "C" int vxw_generate_dense_key(const vxw_gen_params *p, const vxw_key *k, uint16_t *out) {
    if (!p || !k || !out || k->lod > VXW_MAX_LOD) return VXW_E_ARG;
    Gen g(*p);
    if (k->lod == 0) g.chunk(k->pos[0], k->pos[1], k->pos[2], out);
    else g.coarse(Key{k->pos[0], k->pos[1], k->pos[2], (int32_t)k->lod}, out);
    return VXW_OK;
}

extern "C" int vxw_generate(vxw_world *w, const vxw_gen_params *p) {
    if (!w || !p) return VXW_E_ARG;
    for (int i = 0; i < 3; i++)
        if (p->chunk_max[i] < p->chunk_min[i]) return VXW_E_ARG;
    Gen g(*p);
    std::vector<Key> keys;
    for (int32_t y = p->chunk_min[1]; y < p->chunk_max[1]; y++)
        for (int32_t z = p->chunk_min[2]; z < p->chunk_max[2]; z++)
            for (int32_t x = p->chunk_min[0]; x < p->chunk_max[0]; x++) keys.push_back(Key{x, y, z});
    int nt = p->threads > 0 ? p->threads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = std::min<int>(nt, 64);
    // Bounded pipeline: workers build trees for a window of chunks, then the window
    // is inserted in key order so pool offsets do not depend on thread timing.
    const size_t window = 256;
    std::vector<LocalTree> trees(std::min(window, keys.size()));
    for (size_t base = 0; base < keys.size(); base += window) {
        size_t n = std::min(window, keys.size() - base);
        std::atomic<size_t> next{0};
        auto work = [&]() {
            std::vector<BlockId> dense(VOL);
            for (;;) {
                size_t i = next.fetch_add(1);
                if (i >= n) break;
                const Key &k = keys[base + i];
                g.chunk(k.x, k.y, k.z, dense.data());
                trees[i] = build_local(dense.data());
            }
        };
        std::vector<std::thread> th;
        for (int t = 0; t < nt; t++) th.emplace_back(work);
        for (auto &t : th) t.join();
        for (size_t i = 0; i < n; i++) {
            w->w.insert_local(keys[base + i], trees[i]);
            trees[i] = LocalTree();
        }
    }
    return VXW_OK;
}
