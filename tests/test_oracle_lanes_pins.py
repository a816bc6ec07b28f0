"""The oracle's resolve lanes 1 and 2 (glass and water pixels, oracle/vxresolve.inc)
against float64 closed forms.

Scene: one chunk with a stone floor (top at y = 20), a pool of water over it for x < 32
(surface at y = 26, 6 deep; 2 deep where a stone shelf rises to y = 24 for z >= 44), stone
land for x >= 32 (top at y = 26), and a one-voxel glass pane at y = 36 over the land. Open
sky (chunk light uniform, sky level 15), no fog, sky tint / look / cool off, no scattering,
waves off, a per-level uniform atlas with the mip bias at -6 (so every texture lookup but
the water tint reads level 0's colour A0; the water tint under SPEC_WATER_REPAIR reads
level 3, A3), flat light for the legs (SPEC_FLAT_SECONDARY). Then:

- a leg hit on a lit, open floor has the lane-0 closed form without sky specular:
  F(A0) = A0 (0.5 T d + SUN sun_y d 0.78 + FLOOR ambient), T = (0.6038, 0.7084, 1) 0.6;
- a glass pixel (top face of the pane) is mix(F(A0) A0, sky(m), schlick_glass(c)), the
  transmit leg ending on the land under the pane (sun rays pass glass), m the mirror
  about +y, c = -rd_y, schlick_glass(c) = 0.0426 + 0.9574 (1 - c)^5;
- a water pixel is mix(refr, refl, schlick(c)), schlick(c) = 0.02 + 0.98 (1 - c)^5, with
  body = A3 (0.5 d + ambient), refr = F(A0) tr + body (1 - tr), tr = exp(-EXT absorb D),
  D the distance along the Snell-refracted ray (eta = 1/1.333) from 0.02 below the surface
  to the pool floor, and refl = sky(m) + SUN_TINT pow(max(m . sun, 0), 320) 2.5 d (the
  reflect leg leaves the world);
- with SPEC_WATER_SEC, a pixel whose 2x2 block lies on the flat surface uses the legs of
  the block's nearest pixel, stored as f16 (D of that pixel, F(A0) rounded); pixels of
  blocks that are not share-safe equal the resolve without SPEC_WATER_SEC bit for bit;
- shore foam changes only pixels whose column depth is below 2.5 (the shelf), and some of
  them.
Planted mistakes (VXO_PLANT 44-47) fail these checks.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from tests.test_oracle_resolve_pins import (DH, DZ, FLOOR_RGB, LAYERS, LEVEL_RGB, MIPS, NH, NZ, SIZE, SKY_TINT,
                                            SUN_RGB, _atlas, _faces, _srgb64)
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.world import World, build_view

STONE, WATER, GLASS = 1, 13, 10
W_, H_ = 128, 96
SUN_TINT = np.array([1.0, 0.8469, 0.5647])
WATER_EXT = np.array([0.280, 0.055, 0.020])
FLAT_SECONDARY = 1048576
ABSORB = 0.7


def _dense():
    def fn(x, y, z):
        out = np.where(y < 20, STONE, 0)
        out = np.where((x < 32) & (y >= 20) & (y < 26), WATER, out)
        out = np.where((x < 32) & (z >= 44) & (y >= 20) & (y < 24), STONE, out)
        out = np.where((x >= 32) & (y >= 20) & (y < 26), STONE, out)
        return np.where((x >= 40) & (x < 62) & (z >= 4) & (z < 62) & (y == 36), GLASS, out)
    return H.dense_from(fn)


def _scene(sec=False, foam=False, pos=(30.5, 56.3, 2.5), yaw=78.0, pitch=-42.0):
    world = World()
    world.insert_dense((0, 0, 0), _dense())
    flags = (A.FLAG_SHADOWS | A.FLAG_WATER_REFLECT | A.FLAG_WATER_REFRACT | A.FLAG_GLASS | A.FLAG_WATER_SHADOW_CUT
             | FLAT_SECONDARY)
    cam = H.camera(pos, yaw, pitch, W_, H_, flags, sea=26.0)
    v = build_view(world, cam)
    f = v.frame
    f.ambient, f.mip_bias, f.fog_density, f.fog_falloff, f.water_absorb = 0.05, -6.0, 0.0, 0.0, ABSORB
    spec_hi = A.FLAG_HI_FULL_MARCH | (32 if sec else 0) | (16 if foam else 0)
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
    vis, sun, _, _ = O.render(sc, v.frame_bytes, O.spec_from(flags, spec_hi),
                              np.zeros(v.tiles[0] * v.tiles[1], np.uint32))
    out = O.resolve(sc, v.frame_bytes, O.rspec_from(flags, spec_hi), _atlas(), SIZE, LAYERS, MIPS, _faces(), vis, sun)
    return v, vis, out.reshape(-1, 4)


def _rd64(f):
    fwd, right, up = (np.array(x, np.float64) for x in (f.cam_fwd, f.cam_right, f.cam_up))
    px, py = np.meshgrid(np.arange(W_) + 0.5, np.arange(H_) + 0.5)
    nx = px / W_ * 2.0 - 1.0
    ny = 1.0 - py / H_ * 2.0
    d = fwd + right * (nx * f.tan_half_fov * f.aspect)[..., None] + up * (ny * f.tan_half_fov)[..., None]
    return (d / np.linalg.norm(d, axis=-1, keepdims=True)).reshape(-1, 3)


def _lit_floor(f, albedo):
    d = float(f.daylight)
    sy = float(f.sun_dir[1])
    assert sy > 0.0 and d > 0.01
    light = 0.5 * SKY_TINT * 0.6 * d + SUN_RGB * (sy * d) * 0.78 + FLOOR_RGB * float(f.ambient)
    return albedo * light


def _sky(f, rd):
    d = float(f.daylight)
    hz, zn = NH + (DH - NH) * d, NZ + (DZ - NZ) * d
    assert rd[1] >= 0.0
    return hz + (zn - hz) * rd[1] ** 0.55


def _schlick(c, f0):
    m = min(max(1.0 - c, 0.0), 1.0)
    return f0 + (1.0 - f0) * m ** 5


def _refract(rd, eta):
    n = np.array([0.0, 1.0, 0.0])
    d = float(np.dot(n, rd))
    k = 1.0 - eta * eta * (1.0 - d * d)
    return eta * rd - (eta * d + math.sqrt(k)) * n


def _face(key):
    low = key & 0xFFFFFF
    return (key >> 37) & 7, ((low >> 16) & 63, (low >> 8) & 63, low & 63)


def _pixels(v, vis, want_y, x_range):
    """Top-face hits on voxel row y = want_y with the hit well inside x_range and away from
    voxel borders (f32 hit_t agrees with the float64 plane)."""
    f = v.frame
    rd = _rd64(f)
    cam = np.array(f.cam_pos, np.float64)
    out = []
    for i in range(W_ * H_):
        key = int(vis[i])
        if key == 0:
            continue
        nid, vox = _face(key)
        if nid != 3 or vox[1] != want_y:
            continue
        t = (want_y + 1.0 - cam[1]) / rd[i][1]
        hit = cam + rd[i] * t
        if not (x_range[0] <= hit[0] < x_range[1]):
            continue
        if (int(math.floor(hit[0])), int(math.floor(hit[2]))) != (vox[0], vox[2]):
            continue
        out.append((i, t, hit, rd[i]))
    return out


def _glass_misses():
    v, vis, out = _scene()
    f = v.frame
    a0 = _srgb64(LEVEL_RGB[0])
    through = _lit_floor(f, a0)
    bad = n = 0
    for i, t, hit, rd in _pixels(v, vis, 36, (41.0, 61.0)):
        # the transmit leg lands on the land top (y = 26) away from the pool edge
        ro2 = hit + rd * 0.02
        p2 = ro2 + rd * ((26.0 - ro2[1]) / rd[1])
        if not (33.0 <= p2[0] < 63.0 and 1.0 <= p2[2] < 63.0):
            continue  # near the pool edge, or out of the chunk
        m = rd * np.array([1.0, -1.0, 1.0])
        fr = _schlick(-rd[1], 0.0426)
        want = (through * a0) * (1.0 - fr) + _sky(f, m) * fr
        n += 1
        if np.abs(out[i, :3] - want).max() > 2e-5 * max(1.0, np.abs(want).max()) or out[i, 3] != 1.0:
            bad += 1
    assert n > 300, n
    return bad


def _water_want(f, rd, t, hit, depth):
    a0, a3 = _srgb64(LEVEL_RGB[0]), _srgb64(LEVEL_RGB[3])
    d = float(f.daylight)
    body = a3 * (0.5 * d + float(f.ambient))
    tr = np.exp(-WATER_EXT * ABSORB * depth)
    refr = _lit_floor(f, a0) * tr + body * (1.0 - tr)
    m = rd * np.array([1.0, -1.0, 1.0])
    sun = np.array(f.sun_dir, np.float64)
    refl = _sky(f, m) + SUN_TINT * (max(float(np.dot(m, sun)), 0.0) ** 320.0 * 2.5 * d)
    fr = _schlick(-rd[1], 0.02)
    return refr * (1.0 - fr) + refl * fr


def _water_depth(hit, rd):
    """Distance along the refracted leg from 0.02 below the surface to the pool floor
    (y = 20), or None near the pool walls and the shelf."""
    rdir = _refract(rd, 1.0 / 1.333)
    ro2 = hit + rdir * 0.02
    dist = (20.0 - ro2[1]) / rdir[1]
    p2 = ro2 + rdir * dist
    if not (1.0 <= p2[0] < 31.0 and 1.0 <= p2[2] < 43.0 and 1.0 <= hit[2] < 43.0):
        return None
    return dist


def _water_misses():
    v, vis, out = _scene()
    f = v.frame
    bad = n = 0
    for i, t, hit, rd in _pixels(v, vis, 25, (1.0, 31.0)):
        depth = _water_depth(hit, rd)
        if depth is None:
            continue
        want = _water_want(f, rd, t, hit, depth)
        n += 1
        if np.abs(out[i, :3] - want).max() > 2e-5 * max(1.0, np.abs(want).max()) or out[i, 3] != 1.0:
            bad += 1
    assert n > 300, n
    return bad


def test_glass_pixels_closed_form():
    assert _glass_misses() == 0


def test_water_pixels_closed_form():
    assert _water_misses() == 0


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float64)


def test_water_sec_shares_the_block_legs():
    v, vis, sec = _scene(sec=True)
    _, _, per_pixel = _scene(sec=False)
    f = v.frame
    rd_all = _rd64(f)
    cam = np.array(f.cam_pos, np.float64)
    a0 = _srgb64(LEVEL_RGB[0])
    keys = vis.reshape(H_, W_)
    shared = fallback = exact = 0
    for i, t, hit, rd in _pixels(v, vis, 25, (1.0, 31.0)):
        px, py = i % W_, i // W_
        block = [(bx, by) for by in (py & ~1, (py & ~1) + 1) for bx in (px & ~1, (px & ~1) + 1)
                 if bx < W_ and by < H_]
        safe, edge = True, False
        tol = max(0.1, min(1.0, t * 0.02))
        for bx, by in block:
            k = int(keys[by, bx])
            if k == 0 or _face(k)[0] != 3 or _face(k)[1][1] != 25:
                safe = False
                continue
            td = (26.0 - cam[1]) / rd_all[by * W_ + bx][1]
            safe &= abs(td - t) <= tol
            edge |= abs(abs(td - t) - tol) < 1e-3
        if edge:
            continue  # the distance test too close to call at f32 precision
        if not safe:
            assert np.array_equal(sec[i], per_pixel[i])  # shade_water's own legs
            fallback += 1
            continue
        # the block's nearest pixel (x-major order, strict <) gives the legs
        best = None
        for bx in sorted({b[0] for b in block}):
            for by in sorted({b[1] for b in block}):
                rdb = rd_all[by * W_ + bx]
                tb = (26.0 - cam[1]) / rdb[1]
                if best is None or tb < best[0] - 1e-9:
                    best = (tb, cam + rdb * tb, rdb)
        depth = _water_depth(best[1], best[2])
        if depth is None or _water_depth(hit, rd) is None:
            continue
        d = float(f.daylight)
        a3 = _srgb64(LEVEL_RGB[3])
        body = a3 * (0.5 * d + float(f.ambient))
        # the stored depth is the f16 rounding of the f32 leg distance: the float64 value's
        # rounding or, at a rounding boundary, its f16 neighbour
        dep16 = _f16(depth)
        ulp = 2.0 ** (math.floor(math.log2(dep16)) - 10)
        under16 = _f16(_lit_floor(f, a0))
        m = rd * np.array([1.0, -1.0, 1.0])
        sun = np.array(f.sun_dir, np.float64)
        refl = _sky(f, m) + SUN_TINT * (max(float(np.dot(m, sun)), 0.0) ** 320.0 * 2.5 * d)
        fr = _schlick(-rd[1], 0.02)
        err = []
        for dd in (dep16, dep16 - ulp, dep16 + ulp):
            tr = np.exp(-WATER_EXT * ABSORB * dd)
            want = (under16 * tr + body * (1.0 - tr)) * (1.0 - fr) + refl * fr
            err.append(np.abs(sec[i, :3] - want).max() / max(1.0, np.abs(want).max()))
        assert err[0] <= 2e-5 or min(err[1:]) <= 2e-5, (i, sec[i], err)
        exact += err[0] <= 2e-5
        shared += 1
    assert shared > 300 and fallback > 0 and exact > 0.95 * shared, (shared, fallback, exact)


def test_shore_foam_only_in_shallow_water():
    v, vis, foam = _scene(foam=True, pos=(16.5, 50.3, 30.5), yaw=90.0, pitch=-60.0)
    _, _, plain = _scene(foam=False, pos=(16.5, 50.3, 30.5), yaw=90.0, pitch=-60.0)
    changed_shallow = deep = 0
    for i, t, hit, rd in _pixels(v, vis, 25, (1.0, 31.0)):
        rdir = _refract(rd, 1.0 / 1.333)
        ro2 = hit + rdir * 0.02
        p_shelf = ro2 + rdir * ((24.0 - ro2[1]) / rdir[1])
        p_floor = ro2 + rdir * ((20.0 - ro2[1]) / rdir[1])
        if 45.0 <= p_shelf[2] < 62.0 and 1.0 <= p_shelf[0] < 31.0:
            changed_shallow += not np.array_equal(foam[i], plain[i])  # column depth ~2 / |rdir_y|
        elif 1.0 <= p_floor[2] < 43.0 and 1.0 <= p_floor[0] < 31.0 and hit[2] < 43.0:
            assert np.array_equal(foam[i], plain[i])  # column depth ~6: no foam
            deep += 1
    assert changed_shallow > 20 and deep > 100, (changed_shallow, deep)


@pytest.mark.parametrize("plant,check", [
    (44, _water_misses),   # schlick with (1 - c)^4
    (45, _water_misses),   # refraction with eta 1.333 instead of 1 / 1.333
    (46, _glass_misses),   # glass tint not applied to the transmitted colour
    (47, _water_misses),   # water_medium weights swapped
])
def test_planted_lane_mistakes_are_caught(plant, check, tmp_path):
    with O.planted(plant, tmp_path):
        assert check() > 0
