"""The oracle's resolve lane 0 (oracle/vxresolve.inc) against float64 closed forms.

Scene: one chunk with a stone floor (top at y = 20) under open sky (chunk light uniform,
sky level 15: light_curve(15) = 1) and a stone pillar that casts a shadow on it. With
SKY_SPECULAR off, LIGHT_RGB off, no fog and the sky tint off, the lane-0 colour of a floor
pixel is albedo x light with
  light = 0.5 * T * d * ao + SUN * (lit * sun_y * d) * 0.78 + FLOOR * ambient
          (LIGHTING_REPAIR; T = (0.6038, 0.7084, 1) * 0.6), or
  light = (0.5 * T * d + SUN * (lit * sun_y * d) * 0.52) * ao + FLOOR * ambient
          (legacy lighting; T unscaled),
d = daylight, lit = 0 where the shadow word holds a blocker distance and 1 elsewhere,
ao = 1 where the 3x3 cells in front of the
face are air, and for the hand-worked cells beside the pillar: one diagonal cell solid
takes 0.25 off that corner (ao = 1 - 0.25 (1 - fx)(1 - fz) ... by corner), a solid edge
row takes 0.5 off both of its corners. The albedo is the atlas texel: with a per-level
uniform atlas, the linear (sRGB-decoded in float64) colour of level floor(lod) mixed with
the next by fract(lod), lod = clamp(log2(t 2 tan / H 16 / max(|rd_y|, 0.25)) + mip_bias,
0, 3); with a checker on level 0 and lod <= 0, the texel at (floor(16 fx), floor(16 fz)).
Sky pixels (no cloud deck, no scattering, day mix): horizon H = mix(NH, DH, d), zenith
Z = mix(NZ, DZ, d), colour mix(H, Z, up^0.55) above the horizon and
H (1 - 0.74 w), w = smoothstep(0, -0.35, rd_y), below it. Fog (falloff 0): the opaque
colour times exp(-density t) plus the sky times the rest. Heatmap frames: heat(v / 300),
pow(x, 2.2) of the WGSL ramp. log2 / exp2 / pow are bounded against float64.
Planted mistakes (VXO_PLANT 11-15) fail these checks.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.world import World, build_view

STONE = 1
W_, H_ = 96, 64
SIZE, LAYERS, MIPS = 16, 27, 4
SUN_RGB = np.array([1.0, 0.8469, 0.5647])
FLOOR_RGB = np.array([0.85, 0.90, 1.00])
SKY_TINT = np.array([0.6038, 0.7084, 1.0])
NH, DH = np.array([0.0049, 0.0060, 0.0134]), np.array([0.4480, 0.6210, 0.8900])
NZ, DZ = np.array([0.0008, 0.0015, 0.0039]), np.array([0.0637, 0.2140, 0.8276])
LEVEL_RGB = [(200, 40, 10), (30, 180, 60), (20, 50, 220), (240, 240, 30)]


def _srgb64(u):
    c = np.asarray(u, np.float64) / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def _atlas(checker=False):
    levels = []
    for l in range(MIPS):
        d = SIZE >> l
        lv = np.zeros((LAYERS, d, d, 4), np.uint8)
        lv[..., :3] = LEVEL_RGB[l]
        lv[..., 3] = 255
        if checker and l == 0:
            yy, xx = np.meshgrid(np.arange(d), np.arange(d), indexing="ij")
            lv[:, :, :, 0] = (xx * 16 + 7) % 256
            lv[:, :, :, 1] = (yy * 16 + 3) % 256
            lv[:, :, :, 2] = 90
        levels.append(lv.reshape(-1))
    return np.concatenate(levels)


def _faces():
    f = np.zeros((32, 8), np.uint32)
    for i in range(32):
        for n in range(8):
            f[i, n] = (i * 3 + n) % LAYERS
    return f


def _scene(legacy=False, fog=0.0, mip_bias=-6.0, heat=False):
    def fn(x, y, z):
        out = np.where(y < 20, STONE, 0)
        return np.where((x >= 30) & (x < 34) & (z >= 30) & (z < 34) & (y >= 20) & (y < 40), STONE, out)
    dense = H.dense_from(fn)
    world = World()
    world.insert_dense((0, 0, 0), dense)
    flags = A.FLAG_SHADOWS | 8 | (2 if heat else 0)
    cam = H.camera((32.5, 40.3, 2.5), 90.0, -24.0, W_, H_, flags, sea=10.0)
    v = build_view(world, cam)
    v.frame.ambient = 0.05
    v.frame.mip_bias = mip_bias
    v.frame.fog_density = fog
    v.frame.fog_falloff = 0.0
    spec_hi = A.FLAG_HI_FULL_MARCH | (1048576 if legacy else 0)
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
    vis, sun, _, _ = O.render(sc, v.frame_bytes, O.spec_from(flags, spec_hi),
                              np.zeros(v.tiles[0] * v.tiles[1], np.uint32))
    return dense.reshape(64, 64, 64), v, sc, vis, sun, flags, spec_hi


def _rd64(f):
    fwd, right, up = (np.array(x, np.float64) for x in (f.cam_fwd, f.cam_right, f.cam_up))
    px, py = np.meshgrid(np.arange(W_) + 0.5, np.arange(H_) + 0.5)
    nx = px / W_ * 2.0 - 1.0
    ny = 1.0 - py / H_ * 2.0
    d = fwd + right * (nx * f.tan_half_fov * f.aspect)[..., None] + up * (ny * f.tan_half_fov)[..., None]
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def _resolve(sc, v, flags, spec_hi, vis, sun, atlas, heat=None):
    return O.resolve(sc, v.frame_bytes, O.rspec_from(flags, spec_hi), atlas, SIZE, LAYERS, MIPS, _faces(), vis, sun,
                     heat)


def _ao(vx, vz, fx, fz):
    """Hand-worked AO of the floor voxels beside the pillar (x, z in [30, 34))."""
    if (vx, vz) == (29, 29):    # diagonal (+1, +1) solid: corner 3
        return 1.0 - 0.25 * fx * fz
    if (vx, vz) == (34, 29):    # diagonal (-1, +1) solid: corner 2
        return 1.0 - 0.25 * fz * (1.0 - fx)
    if vx == 29 and 31 <= vz <= 32:  # the +x row solid: corners 1 and 3 at 0.5
        return 1.0 - 0.5 * fx
    return None


def _floor_pixels(dense, v, vis):
    """(pixel, t, hit, voxel) of floor-top hits, from float64 geometry, checked against the key."""
    f = v.frame
    rd = _rd64(f).reshape(-1, 3)
    cam = np.array(f.cam_pos, np.float64)
    out = []
    for i in range(W_ * H_):
        key = int(vis[i])
        if key == 0 or ((key >> 37) & 7) != 3:
            continue
        low = key & 0xFFFFFF
        vox = ((low >> 16) & 63, (low >> 8) & 63, low & 63)
        if vox[1] != 19:
            continue
        t = (20.0 - cam[1]) / rd[i][1]
        hit = cam + rd[i] * t
        if (int(math.floor(hit[0])), int(math.floor(hit[2]))) != (vox[0], vox[2]):
            continue  # on a voxel border at f32 precision
        out.append((i, t, hit, vox, rd[i]))
    return out


def _open(dense, vx, vz):
    ys = dense[20]
    return not ys[max(vz - 1, 0):vz + 2, max(vx - 1, 0):vx + 2].any()


def _light(v, lit, ao, legacy):
    f = v.frame
    d = float(f.daylight)
    sy = float(f.sun_dir[1])
    direct = lit * sy * d if sy > 0.0 and d > 0.01 else 0.0
    if legacy:
        return (0.5 * SKY_TINT * d + SUN_RGB * direct * 0.52) * ao + FLOOR_RGB * float(f.ambient)
    return 0.5 * SKY_TINT * 0.6 * d * ao + SUN_RGB * direct * 0.78 + FLOOR_RGB * float(f.ambient)


def _lod_raw(v, t, rdy):
    f = v.frame
    fp = max(t, 1e-3) * 2.0 * f.tan_half_fov / H_ * 16.0 / max(abs(rdy), 0.25)
    return math.log2(max(fp, 1e-4)) + f.mip_bias


def _lod(v, t, rdy):
    return min(max(_lod_raw(v, t, rdy), 0.0), 3.0)


def _floor_misses(legacy):
    dense, v, sc, vis, sun, flags, spec_hi = _scene(legacy)
    out = _resolve(sc, v, flags, spec_hi, vis, sun, _atlas()).reshape(-1, 4)
    lin = [_srgb64(c) for c in LEVEL_RGB]
    bad = n = shadowed = worked = 0
    for i, t, hit, vox, rd in _floor_pixels(dense, v, vis):
        fx, fz = hit[0] - vox[0], hit[2] - vox[2]
        ao = 1.0 if _open(dense, vox[0], vox[2]) else _ao(vox[0], vox[2], fx, fz)
        if ao is None:
            continue
        worked += ao < 1.0
        lit = 0.0 if float(np.array([sun[i]], np.uint16).view(np.float16)[0]) > 0 else 1.0
        shadowed += lit == 0.0
        lod = _lod(v, t, rd[1])
        l0 = int(math.floor(lod))
        l1 = min(l0 + 1, MIPS - 1)
        albedo = lin[l0] * (1.0 - (lod - l0)) + lin[l1] * (lod - l0)
        want = albedo * _light(v, lit, ao, legacy)
        n += 1
        if np.abs(out[i, :3] - want).max() > 2e-6 * max(1.0, np.abs(want).max()) or out[i, 3] != 1.0:
            bad += 1
    assert n > 500 and shadowed > 50 and worked >= 5, (n, shadowed, worked)
    return bad


@pytest.mark.parametrize("legacy", [False, True])
def test_floor_pixels_closed_form(legacy):
    assert _floor_misses(legacy) == 0


def _texel_misses():
    dense, v, sc, vis, sun, flags, spec_hi = _scene(mip_bias=-12.0)
    atlas = _atlas(checker=True)
    out = _resolve(sc, v, flags, spec_hi, vis, sun, atlas).reshape(-1, 4)
    bad = n = 0
    for i, t, hit, vox, rd in _floor_pixels(dense, v, vis):
        if not _open(dense, vox[0], vox[2]) or _lod_raw(v, t, rd[1]) > -1e-3:
            continue
        fx, fz = hit[0] - vox[0], hit[2] - vox[2]
        if min(fx * 16 % 1, fz * 16 % 1) < 1e-3 or max(fx * 16 % 1, fz * 16 % 1) > 1 - 1e-3:
            continue  # on a texel border at f32 precision
        tx, tz = int(fx * 16), int(fz * 16)  # top face: u = local x, v = local z
        texel = np.array([(tx * 16 + 7) % 256, (tz * 16 + 3) % 256, 90])
        lit = 0.0 if float(np.array([sun[i]], np.uint16).view(np.float16)[0]) > 0 else 1.0
        want = _srgb64(texel) * _light(v, lit, 1.0, False)
        n += 1
        if np.abs(out[i, :3] - want).max() > 2e-6:
            bad += 1
    assert n > 300, n
    return bad


def test_nearest_texel_at_lod_zero():
    assert _texel_misses() == 0


def _sky_misses():
    dense, v, sc, vis, sun, flags, spec_hi = _scene()
    out = _resolve(sc, v, flags, spec_hi, vis, sun, _atlas()).reshape(-1, 4)
    f = v.frame
    d = float(f.daylight)
    hz, zn = NH + (DH - NH) * d, NZ + (DZ - NZ) * d
    sun_d = np.array(f.sun_dir, np.float64)
    rd = _rd64(f).reshape(-1, 3)
    bad = n = below = 0
    for i in range(W_ * H_):
        if vis[i] != 0 or np.dot(rd[i], sun_d) > 0.997:
            continue
        y = rd[i][1]
        if y >= 0:
            want = hz + (zn - hz) * y ** 0.55
        else:
            s = min(max(y / -0.35, 0.0), 1.0)
            w = s * s * (3 - 2 * s)
            want = hz * (1 - w) + hz * 0.26 * w
            below += 1
        n += 1
        if np.abs(out[i, :3] - want).max() > 3e-6:
            bad += 1
    assert n > 300 and below > 20, (n, below)
    return bad


def test_sky_pixels_closed_form():
    assert _sky_misses() == 0


def test_fog_mixes_toward_the_sky():
    dense, v, sc, vis, sun, flags, spec_hi = _scene(fog=0.01)
    _, v0, sc0, vis0, sun0, _, _ = _scene(fog=0.0)
    fogged = _resolve(sc, v, flags, spec_hi, vis, sun, _atlas()).reshape(-1, 4)
    clear = _resolve(sc0, v0, flags, spec_hi, vis0, sun0, _atlas()).reshape(-1, 4)
    f = v.frame
    d = float(f.daylight)
    hz, zn = NH + (DH - NH) * d, NZ + (DZ - NZ) * d
    n = 0
    for i, t, hit, vox, rd in _floor_pixels(dense, v, vis):
        tr = math.exp(-0.01 * t)
        y = rd[1]
        s = min(max(y / -0.35, 0.0), 1.0)
        w = s * s * (3 - 2 * s)
        sky = hz * (1 - w) + hz * 0.26 * w if y < 0 else hz + (zn - hz) * y ** 0.55
        want = clear[i, :3] * tr + sky * (1 - tr)
        assert np.abs(fogged[i, :3] - want).max() <= 3e-6, (i, fogged[i], want)
        n += 1
    assert n > 500


def test_heatmap_frame_closed_form():
    dense, v, sc, vis, sun, flags, spec_hi = _scene(heat=True)
    heat = (np.arange(W_ * H_, dtype=np.uint32) * 7) % 700
    out = _resolve(sc, v, flags, spec_hi, vis, sun, _atlas(), heat=heat).reshape(-1, 4)

    def sm(a, b, x):
        t = np.clip((x - a) / (b - a), 0.0, 1.0)
        return t * t * (3 - 2 * t)
    t = np.clip(heat / 300.0, 0.0, 1.0)
    c = np.stack([sm(0.35, 0.85, t), sm(0.0, 0.45, t) - sm(0.65, 1.0, t), 1.0 - sm(0.0, 0.45, t)], axis=-1)
    want = np.where(c > 0, np.abs(c) ** 2.2, 0.0)
    assert np.abs(out[:, :3] - want).max() <= 2e-6
    assert (out[:, 3] == 1.0).all()


def test_elementary_functions_bounds():
    L = O.lib()
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(1e-30, 1.0, 3000), rng.uniform(1.0, 1e30, 3000), [1.0, 2.0, 0.5, 1e-40, 3e-39],
                         np.exp(rng.uniform(-80, 80, 3000))]).astype(np.float32)
    for x in xs:
        got = L.vxo_log2(float(x))
        want = math.log2(float(x))
        assert abs(got - want) <= 3e-7 * max(1.0, abs(want)), (x, got, want)
    ys = rng.uniform(-140, 120, 6000).astype(np.float32)
    for y in ys:
        got, want = L.vxo_exp2(float(y)), 2.0 ** float(y)
        if want >= 2.0 ** -126:
            assert abs(got - want) <= 2 * 2.0 ** -23 * want, (y, got, want)
        else:
            assert abs(got - want) <= 2.0 ** -149, (y, got, want)
    assert L.vxo_log2(0.0) == -math.inf and L.vxo_exp2(-200.0) == 0.0 and L.vxo_pow(0.0, 2.2) == 0.0
    for _ in range(4000):
        x = float(np.float32(rng.uniform(0.0, 4.0)))
        y = float(np.float32(rng.choice([0.55, 1.5, 2.2, 2.4, 2.5, 3.0, 4.0])))
        got, want = L.vxo_pow(x, y), x ** y
        e = abs(y * math.log2(x)) if x > 0 else 0.0
        assert abs(got - want) <= 2.5e-7 * (1.0 + e) * want + 1e-38, (x, y, got, want)


@pytest.mark.parametrize("plant,check", [
    (11, lambda: _floor_misses(False)),   # the sun word ignored
    (12, lambda: _texel_misses()),        # top-face uv axes swapped
    (13, lambda: _sky_misses()),          # below-horizon mix toward the zenith
    (14, lambda: _floor_misses(False)),   # AO corner without its diagonal cell
    (15, lambda: _floor_misses(True)),    # texels not decoded from sRGB
])
def test_planted_resolve_mistakes_are_caught(plant, check, tmp_path):
    with O.planted(plant, tmp_path):
        assert check()
