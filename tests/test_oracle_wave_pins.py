"""The oracle's wave field and water reflect direction against float64 mathematics.

What is pinned (resolve.wgsl:902-967 for the field, :1078-1167 for water_ctx's mirror_t):
- cos and exp: WGSL leaves their accuracy to the implementation; the oracle's f32 versions
  (shared with the HIP kernels) are bounded against float64 on the whole range the wave
  field feeds them: |cos error| <= 2e-7, exp within one ulp;
- the field's slope gradient: the octaves are plane waves of height
  s_i f_i lambda_i / (2 pi) sin(2 pi / lambda_i <p, d_i> + phase_i), so the field must equal
  the gradient of their float64 sum, taken by central differences (tables transcribed
  from resolve.wgsl:914-927 and :950-966, at a non-zero animation time so the phase
  counts); independently of any table, the field's Jacobian must be symmetric (a
  gradient has no curl);
- the unresolved variance: 0.5 sum s_i^2 (1 - f_i^2), with f_i = 0 and 1 at the two ends
  of the isotropic filter and f_i = exp(-(px sqrt(1 + kk <d_i, vh>^2) / lambda_i)^2) with
  the anisotropic one;
- the reflect direction: the ray mirrored about normalize(n_t (1 - r) + n r), with
  n_t = normalize(-g_x m, 1, -g_z m), m = min(1, slope / |g|), r = clamp(1 - exp(-8 u)),
  then y clamped to >= 0 and renormalised; the shoal factor smoothstep(0, 6, depth) from
  the distance straight down to the floor of a shallow pool; the flat mirror when waves
  are off or the face is not a top face.
Copies of the oracle built with a planted mistake (VXO_PLANT 7-10 in oracle/vxoracle.c:
phase sign, swapped direction components, exp(-u) for exp(-u^2), roughness dropped) fail
these checks.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.world import World, build_view

STONE, WATER = 1, 13
FILL = 4096

# (direction, wavelength divisor, amplitude, phase multiplier), resolve.wgsl:914-927, :950-966
OCT_D = [((0.913089, 0.407760), 1.0, 0.42, 1.0000), ((-0.388685, 0.921371), 2.17, 0.27, 1.4731),
         ((-0.980382, -0.197108), 4.71, 0.19, 2.1703), ((0.477328, -0.878725), 10.20, 0.12, 3.1937)]
OCT_F = [((-0.438371, 0.898794), 1.0, 0.3086, 1.0000), ((-0.927184, 0.374607), 1.3934, 0.2584, 1.1804),
         ((-0.190809, 0.981627), 1.9417, 0.2163, 1.3934), ((-0.669131, 0.743145), 2.7056, 0.1810, 1.6449),
         ((0.121869, 0.992546), 3.7700, 0.1515, 1.9417), ((0.694658, 0.719340), 5.2533, 0.1269, 2.2920),
         ((0.970296, 0.241922), 7.3201, 0.1062, 2.7056), ((0.438371, 0.898794), 10.2000, 0.0889, 3.1937)]
AMP, SCALE, SPEED, TIME = 0.12, 32.0, 0.32, 1.7


def _frame(fill=False, amp=AMP, scale=SCALE, speed=SPEED, time=TIME, slope=0.0, res_y=64, tan_half=0.7):
    f = A.Frame()
    f.flags = 512 | 1024 | 2048 | (FILL if fill else 0)
    f.wave_amp, f.wave_scale, f.wave_speed, f.wave_reflect_slope = amp, scale, speed, slope
    f.time = time
    f.res[0], f.res[1] = res_y * 3 // 2, res_y
    f.tan_half_fov = tan_half
    return f


def _spec(fill=False):
    return O.spec_from(512 | 1024 | 2048 | (FILL if fill else 0), 0)


def _field(f, s, x, z, px, vh=(0.0, 0.0), kk=0.0, aniso=False, shoal=1.0):
    out = (C.c_float * 3)()
    O.lib().vxo_wave_field(C.byref(f), C.byref(s), x, z, px, (C.c_float * 2)(*vh), kk, int(aniso), shoal, out)
    return np.array(out[:], np.float64)


def _octaves(fill):
    return OCT_F if fill else OCT_D


def _height64(x, z, fill, fades, amp=AMP, scale=SCALE, speed=SPEED, time=TIME, shoal=1.0):
    h = 0.0
    for (d, div, a, phm), fd in zip(_octaves(fill), fades):
        lam = scale / div
        s = amp * shoal * a
        ph = time * speed * 2.0 * math.pi * phm
        h += s * fd * lam / (2.0 * math.pi) * math.sin(2.0 * math.pi / lam * (x * d[0] + z * d[1]) + ph)
    return h


def _grad64(x, z, fill, fades, **kw):
    e = 1e-5
    gx = (_height64(x + e, z, fill, fades, **kw) - _height64(x - e, z, fill, fades, **kw)) / (2 * e)
    gz = (_height64(x, z + e, fill, fades, **kw) - _height64(x, z - e, fill, fades, **kw)) / (2 * e)
    return np.array([gx, gz])


def _points(n=300, seed=5, span=300.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(-span, span, (n, 2)).astype(np.float32)


# --- cos / exp -------------------------------------------------------------------------

def test_wave_cos_bound():
    L = O.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-8000, 8000, 40000), rng.uniform(-8, 8, 10000),
                         np.arange(-64, 65) * (math.pi / 4)]).astype(np.float32)
    got = np.array([L.vxo_wave_cos(float(x)) for x in xs])
    err = np.abs(got - np.cos(xs.astype(np.float64)))
    assert err.max() <= 2e-7, (err.max(), xs[err.argmax()])
    assert L.vxo_wave_cos(0.0) == 1.0


def test_wave_exp_bound():
    L = O.lib()
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(-86, 20, 40000), [-86.0, 0.0, 1.0, -1.0]]).astype(np.float32)
    got = np.array([L.vxo_wave_exp(float(x)) for x in xs])
    ref = np.exp(xs.astype(np.float64))
    ulp = np.abs(got - ref) / (ref * 2.0 ** -23)
    assert ulp.max() <= 1.0, (ulp.max(), xs[ulp.argmax()])
    assert L.vxo_wave_exp(-86.5) == 0.0 and L.vxo_wave_exp(-1000.0) == 0.0
    assert L.vxo_wave_exp(0.0) == 1.0


# --- the field --------------------------------------------------------------------------

def _gradient_misses(fill):
    """Points where the field (fade 1 everywhere: px tiny, isotropic) is not the gradient
    of the float64 height sum."""
    f, s = _frame(fill), _spec(fill)
    n = len(_octaves(fill))
    bad = 0
    for x, z in _points():
        got = _field(f, s, float(x), float(z), 1e-3)
        want = _grad64(float(x), float(z), fill, [1.0] * n)
        if np.abs(got[:2] - want).max() > 2e-6 or got[2] != 0.0:
            bad += 1
    return bad


@pytest.mark.parametrize("fill", [False, True])
def test_field_is_the_gradient_of_the_height_sum(fill):
    assert _gradient_misses(fill) == 0


def _curl(fill):
    """Largest |d gx / dz - d gz / dx| of the field by central differences in f32 steps,
    relative to the largest derivative seen."""
    f, s = _frame(fill), _spec(fill)
    h = 1.0 / 64.0
    worst, scale = 0.0, 0.0
    for x, z in _points(120, seed=9, span=200.0):
        x, z = float(np.float32(x)), float(np.float32(z))
        gx_dz = (_field(f, s, x, z + h, 1e-3)[0] - _field(f, s, x, z - h, 1e-3)[0]) / (2 * h)
        gz_dx = (_field(f, s, x + h, z, 1e-3)[1] - _field(f, s, x - h, z, 1e-3)[1]) / (2 * h)
        worst = max(worst, abs(gx_dz - gz_dx))
        scale = max(scale, abs(gx_dz), abs(gz_dx))
    return worst / scale


@pytest.mark.parametrize("fill", [False, True])
def test_field_has_no_curl(fill):
    assert _curl(fill) < 2e-3


@pytest.mark.parametrize("fill", [False, True])
def test_unresolved_variance_isotropic_ends(fill):
    f, s = _frame(fill), _spec(fill)
    full = 0.5 * sum((AMP * a) ** 2 for _, _, a, _ in _octaves(fill))
    for x, z in _points(20):
        far = _field(f, s, float(x), float(z), 1e6)
        assert far[0] == 0.0 and far[1] == 0.0
        assert abs(far[2] - full) <= 1e-6 * full
        assert _field(f, s, float(x), float(z), 1e-3)[2] == 0.0
    # half shoal: amplitudes halve, so the variance quarters
    assert abs(_field(f, s, 1.0, 2.0, 1e6, shoal=0.5)[2] - full / 4) <= 1e-6 * full


def _aniso_misses(fill):
    f, s = _frame(fill), _spec(fill)
    rng = np.random.default_rng(11)
    bad = 0
    for _ in range(200):
        ang = rng.uniform(0, 2 * math.pi)
        vh = (math.cos(ang), math.sin(ang))
        vh32 = np.array(vh, np.float32).astype(np.float64)
        kk = float(np.float32(rng.uniform(0.0, 40.0)))
        px = float(np.float32(rng.uniform(0.05, 4.0)))
        x, z = (float(v) for v in rng.uniform(-100, 100, 2).astype(np.float32))
        got = _field(f, s, x, z, px, vh=vh, kk=kk, aniso=True)
        fades, var = [], 0.0
        for d, div, a, _ in _octaves(fill):
            lam = SCALE / div
            dd = d[0] * vh32[0] + d[1] * vh32[1]
            u = px * math.sqrt(1.0 + kk * dd * dd) / lam
            fd = math.exp(-u * u)
            fades.append(fd)
            var += 0.5 * (AMP * a) ** 2 * (1.0 - fd * fd)
        want = _grad64(x, z, fill, fades)
        if abs(got[2] - var) > 2e-5 * max(var, 1e-3) or np.abs(got[:2] - want).max() > 2e-6:
            bad += 1
    return bad


@pytest.mark.parametrize("fill", [False, True])
def test_anisotropic_fade_closed_form(fill):
    assert _aniso_misses(fill) == 0


# --- the reflect direction --------------------------------------------------------------

def _pool_world():
    """Deep water (12 blocks) over a stone floor, with a shallow corner (3 blocks) and a
    dry island."""
    def fn(x, y, z):
        out = np.where(y < 20, STONE, np.where(y < 32, WATER, 0))
        out = np.where((x < 16) & (z < 16) & (y < 29), STONE, out)
        out = np.where((x >= 40) & (x < 44) & (z >= 40) & (z < 44) & (y < 34), STONE, out)
        return out
    world = World()
    world.insert_dense((0, 0, 0), H.dense_from(fn))
    return world


def _reflect_setup(flags, slope, fill=False):
    world = _pool_world()
    cam = H.camera((30.5, 60.0, 30.5), 45.0, -30.0, 96, 64, flags, sea=32.0)
    cam.anim_time = TIME
    cam.waves = (AMP, SCALE, SPEED, slope)
    v = build_view(world, cam)
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
    return sc, v


def _reflect(sc, v, sp, nid, hit, rd, t):
    f = np.frombuffer(v.frame_bytes, np.uint8).copy()
    out = (C.c_float * 3)()
    O.lib().vxo_water_reflect_dir(C.byref(sc.pools), O._p(sc.chunks), sc.n_chunks, O._p(sc.grid), O._p(f),
                                  C.byref(sp), nid, (C.c_float * 3)(*hit), (C.c_float * 3)(*rd), t, out)
    return np.array(out[:], np.float64)


def _smooth(e0, e1, x):
    s = min(max((x - e0) / (e1 - e0), 0.0), 1.0)
    return s * s * (3.0 - 2.0 * s)


def _reflect64(v, hit, rd, t, slope, aniso, shoal_on, depth, fill=False):
    fr = v.frame
    rd = np.asarray(rd, np.float64)
    px = max(t, 1e-3) * 2.0 * fr.tan_half_fov / fr.res[1]
    kk, vh = 0.0, np.zeros(2)
    if aniso:
        k = 1.0 / max(abs(rd[1]), 1e-3)
        kk = k * k - 1.0
        vh = np.array([rd[0], rd[2]]) / math.sqrt(max(rd[0] ** 2 + rd[2] ** 2, 1e-12))
    shoal = _smooth(0.0, 6.0, depth) if shoal_on else 1.0
    fades, var = [], 0.0
    for d, div, a, _ in _octaves(fill):
        lam = SCALE / div
        if aniso:
            dd = d[0] * vh[0] + d[1] * vh[1]
            u = px * math.sqrt(1.0 + kk * dd * dd) / lam
            fd = math.exp(-u * u)
        else:
            fd = _smooth(2.0, 4.0, lam / max(px, 1e-4))
        fades.append(fd)
        var += 0.5 * (AMP * shoal * a) ** 2 * (1.0 - fd * fd)
    g = _grad64(hit[0], hit[2], fill, fades, shoal=shoal)
    m = min(1.0, slope / max(np.linalg.norm(g), 1e-6))
    nt = np.array([-g[0] * m, 1.0, -g[1] * m])
    nt /= np.linalg.norm(nt)
    n = np.array([0.0, 1.0, 0.0])
    r = min(max(1.0 - math.exp(-8.0 * max(var, 0.0)), 0.0), 1.0)
    nt = nt * (1.0 - r) + n * r
    nt /= np.linalg.norm(nt)
    out = rd - 2.0 * np.dot(nt, rd) * nt
    out[1] = max(out[1], 0.0)
    return out / np.linalg.norm(out)


def _reflect_misses(slope, aniso, shoal_on, fill=False):
    flags = 512 | (1024 if aniso else 0) | (2048 if shoal_on else 0) | (FILL if fill else 0)
    sc, v = _reflect_setup(flags, slope, fill)
    sp = O.spec_from(flags, 0)
    rng = np.random.default_rng(21)
    bad, rough = 0, 0
    for i in range(160):
        corner = i % 2 == 0  # half the points over the 3-block shallows
        x = float(np.float32(rng.uniform(1.0, 15.0) if corner else rng.uniform(17.0, 39.0)))
        z = float(np.float32(rng.uniform(1.0, 15.0) if corner else rng.uniform(17.0, 39.0)))
        d = np.array([rng.uniform(-1, 1), -rng.uniform(0.05, 1.0), rng.uniform(-1, 1)])
        d = (d / np.linalg.norm(d)).astype(np.float32)
        t = float(np.float32(rng.uniform(2.0, 400.0)))
        hit = np.array([x, 32.0, z], np.float32)
        got = _reflect(sc, v, sp, 3, hit, d, t)
        # the shoal probe starts 0.02 below the surface and meets the floor below it
        depth = (32.0 - 0.02 - (29.0 if corner else 20.0)) / 0.999999
        depth = min(depth, 6.0)
        want = _reflect64(v, hit.astype(np.float64), d, t, slope, aniso, shoal_on, depth, fill)
        if np.abs(got - want).max() > 2e-5:
            bad += 1
        if abs(np.dot(want, [0, 1, 0]) + d[1]) > 1e-3:
            rough += 1  # not the flat mirror: the wave normal mattered
    assert rough > 40
    return bad


@pytest.mark.parametrize("slope,aniso,shoal,fill", [(0.35, True, True, False), (0.35, False, True, True),
                                                   (0.2, True, False, False), (1.0, True, True, True)])
def test_reflect_direction_float64(slope, aniso, shoal, fill):
    assert _reflect_misses(slope, aniso, shoal, fill) == 0


def test_reflect_direction_is_the_flat_mirror_without_waves():
    sc, v = _reflect_setup(A.FLAG_WATER_REFLECT, 0.35)
    rng = np.random.default_rng(4)
    for spec_flags, nid in [(0, 3), (512 | 1024, 0), (512 | 1024, 5)]:
        sp = O.spec_from(spec_flags, 0)
        for _ in range(20):
            d = np.array([rng.uniform(-1, 1), -rng.uniform(0.05, 1.0), rng.uniform(-1, 1)])
            d = (d / np.linalg.norm(d)).astype(np.float32)
            n = np.zeros(3, np.float32)
            n[nid >> 1] = -1.0 if nid & 1 == 0 else 1.0
            got = _reflect(sc, v, sp, nid, np.array([20.5, 32.0, 20.5], np.float32), d, 50.0)
            want = d.astype(np.float64) - 2.0 * float(np.dot(n, d)) * n
            assert np.abs(got - want).max() <= 1e-6


# --- planted mistakes --------------------------------------------------------------------

@pytest.mark.parametrize("plant,check", [
    (7, lambda: _gradient_misses(False)),                       # phase sign
    (8, lambda: _curl(True) > 2e-3),                            # d.x and d.y swapped in <p, d>
    (9, lambda: _aniso_misses(False)),                          # exp(-u) for exp(-u^2)
    (10, lambda: _reflect_misses(1.0, True, True, True)),       # roughness blend dropped
])
def test_planted_wave_mistakes_are_caught(plant, check, tmp_path):
    with O.planted(plant, tmp_path):
        assert check()
