"""Distant terrain shadows: the shaft height window (vxw_shaft_*, src/shaft.rs), the oracle's
shaft_scan (shaft.wgsl) and terrain_shade_beyond in the resolve (common.wgsl:1557-1580,
resolve.wgsl:516-518).

Host window, ports of tests/shaft.rs on this build's terrain generator: the window scrolls
without changing what it holds, an unmoved camera refills nothing, the brute-force
envelope falls away from the sun and leaves the terrain alone under an overhead sun, and
the envelope from the handoff distance is the offset query terrain_shade_beyond makes.

Oracle scan, pinned by
- tests/shaft.rs's own bound: the scan is the brute-force envelope within SOFT / 2, and
  it raises a quarter of the columns above their ground (under a lower sun than the
  reference's test, for this generator's gentler terrain);
- a closed form: one spike of height H on flat ground, the sun in the x-y plane: the
  scan's result 1 <= m <= 255 texels from the spike, away from the sun, is
  H - m texel tan(elevation), and 0 towards the sun.
terrain_shade, pinned on a result that is a ramp in x (exact in f32): the smoothstep of
p.y around the ramp's value at the texel p falls on (texel centres at origin +
(i + 0.5) texel), at p + sun d for the `beyond` form.
Resolve, on the floor-and-pillar chunk of the lane-0 pins: a field far below changes no
bit, a field far above darkens every sun term as blocked shadow words do, and with a
field whose edge ramp crosses the floor's query point each floor pixel's colour is the
closed form with lit x smoothstep(e - 4, e + 4, 20 + sun_y shadow_dist).
Planted mistakes 48-52 each fail one of these.
"""
import dataclasses
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import test_oracle_resolve_pins as RP
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.world import ShaftField

DIM, SOFT, TEXEL = 512, 4.0, 4
DISTANT = 32768


def _gen(seed):
    return dataclasses.replace(W.WORKLOADS[1].gen, seed=seed)


def _frame(sun, texel=TEXEL, origin=(0.0, 0.0), soft=SOFT):
    f = A.Frame()
    f.sun_dir[0], f.sun_dir[1], f.sun_dir[2] = (float(v) for v in sun)
    f.shaft_texel = float(texel)
    f.shaft_origin[0], f.shaft_origin[1] = origin
    f.shaft_soft = soft
    return f


def _smoothstep64(e0, e1, x):
    t = min(max((x - e0) / (e1 - e0), 0.0), 1.0)
    return t * t * (3.0 - 2.0 * t)


# ---- host window (ports of tests/shaft.rs) ----

def test_the_window_scrolls_without_changing_what_it_holds():
    gen = _gen(11)
    a = ShaftField()
    a.update(gen, 0.0, 0.0)
    shift = float(40 * TEXEL)
    b = ShaftField()
    b.update(gen, shift, shift)
    oa, ob = a.origin_world(), b.origin_world()
    assert (ob[0] - oa[0], ob[1] - oa[1]) == (shift, shift)
    scrolled = ShaftField()
    scrolled.update(gen, 0.0, 0.0)
    scrolled.update(gen, shift, shift)
    assert scrolled.origin_world() == ob
    assert np.array_equal(scrolled.heights(), b.heights())


def test_an_unmoved_camera_refills_nothing():
    gen = _gen(5)
    f = ShaftField()
    f.update(gen, 100.0, -60.0)
    assert f.dirty()
    for _ in range(8):
        f.update(gen, 100.0, -60.0)
        assert not f.dirty()
    f.update(gen, 100.0 + TEXEL * 0.4, -60.0)
    assert not f.dirty()
    with pytest.raises(Exception):
        f.update(gen, 0.0, 0.0, canopy_lift=2.0)  # tree-line canopies are not restated
    with pytest.raises(ValueError):
        ShaftField(texel=65)


def test_the_shadow_falls_away_from_the_sun_and_a_high_sun_leaves_the_terrain_alone():
    f = ShaftField()
    f.update(_gen(7), 0.0, 0.0)
    h = f.heights()
    sun = (1.0, 0.05, 0.0)  # a low sun along +x: 0.2 blocks of climb a texel
    k = 4
    rise = h[:, k:] - h[:, :-k]  # the ground k texels towards the sun, above a column
    row, col = np.unravel_index(np.argmax(rise), rise.shape)
    assert rise[row, col] > k * TEXEL * 0.05 + 2.0, rise[row, col]
    assert f.envelope_reference(sun, col, row) > h[row, col] + 1.0  # below the rise: shadowed
    pz, px = np.unravel_index(np.argmax(h), h.shape)
    assert f.envelope_reference(sun, px, pz) <= h[pz, px] + 1e-3  # the highest column: lit
    for tz in range(0, DIM, 101):
        for tx in range(0, DIM, 101):
            assert abs(f.envelope_reference((0.001, 1.0, 0.0), tx, tz) - h[tz, tx]) < 1e-3


def _envelope_from(h, sun, tx, tz, m0):
    horiz = max(math.hypot(sun[0], sun[2]), 1e-3)
    dx, dz, slope = sun[0] / horiz, sun[2] / horiz, sun[1] / horiz

    def sample(x, z):
        cx, cz = min(max(x, 0.0), DIM - 1.0), min(max(z, 0.0), DIM - 1.0)
        x0, z0 = int(math.floor(cx)), int(math.floor(cz))
        fx, fz = cx - x0, cz - z0
        x1, z1 = min(x0 + 1, DIM - 1), min(z0 + 1, DIM - 1)
        top = h[z0, x0] + (h[z0, x1] - h[z0, x0]) * fx
        bot = h[z1, x0] + (h[z1, x1] - h[z1, x0]) * fx
        return top + (bot - top) * fz

    return max(sample(tx + dx * m, tz + dz * m) - m * TEXEL * slope for m in range(m0, m0 + 257))


def test_the_offset_is_the_envelope_with_the_near_occluders_struck_out():
    f = ShaftField()
    f.update(_gen(1337), 0.0, 0.0)
    h = f.heights().astype(np.float64)
    sun = (0.82, 0.18, 0.54)
    horiz = math.hypot(sun[0], sun[2])
    dx, dz, slope = sun[0] / horiz, sun[2] / horiz, sun[1] / horiz
    hand = 55  # 220 blocks at 4 blocks a texel: the march's shadow_dist
    lift = hand * TEXEL * slope
    worst, checked, lower, above = 0.0, 0, 0, 0
    for tz in range(0, DIM, 37):
        for tx in range(0, DIM, 37):
            want = _envelope_from(h, sun, tx, tz, hand)
            have = _envelope_from(h, sun, tx + dx * hand, tz + dz * hand, 0) - lift
            worst = max(worst, abs(want - have))
            full = _envelope_from(h, sun, tx, tz, 0)
            lower += want < full - 0.5
            above += want > full + 1e-2
            checked += 1
    assert checked > 150 and worst < 1e-2 and above == 0 and lower * 4 > checked, (worst, above, lower, checked)


# ---- oracle scan ----

def _scan_vs_envelope():
    f = ShaftField()
    f.update(_gen(1337), 0.0, 0.0)
    # tests/shaft.rs's sun direction, lower (y 0.04 for 0.18): this generator's terrain is
    # gentler than worldgen.rs's, so the reference's elevation raises few columns
    sun = (0.82, 0.04, 0.54)
    got = O.shaft_scan(f.heights(), bytes(_frame(sun)))
    h = f.heights()
    worst, checked, moved = 0.0, 0, 0
    for tz in range(0, DIM, 37):
        for tx in range(0, DIM, 37):
            worst = max(worst, abs(f.envelope_reference(sun, tx, tz) - got[tz, tx]))
            moved += got[tz, tx] > h[tz, tx] + 0.5
            checked += 1
    assert checked > 150
    assert worst < SOFT * 0.5, worst
    assert moved * 4 > checked, (moved, checked)


def test_oracle_scan_is_the_envelope_it_claims_to_be():
    _scan_vs_envelope()


def _spike():
    h = np.zeros((DIM, DIM), np.float32)
    sx, sz, hh = 300, 256, 200.0
    h[sz, sx] = hh
    elev = math.radians(20.0)
    sun = (math.cos(elev), math.sin(elev), 0.0)
    got = O.shaft_scan(h, bytes(_frame(sun)))
    sunf = np.array(sun, np.float32).astype(np.float64)
    slope = sunf[1] / sunf[0]
    m = np.arange(1, 256)
    want = np.maximum(hh - m * TEXEL * slope, 0.0)
    assert np.allclose(got[sz, sx - m], want, atol=1e-3), np.abs(got[sz, sx - m] - want).max()
    assert got[sz, sx] == hh
    assert np.all(got[sz, sx + 1:] == 0.0) and np.all(got[sz + 2:, :] == 0.0) and np.all(got[:sz - 1, :] == 0.0)


def test_oracle_scan_of_a_spike_is_its_closed_form():
    _spike()


def _ramp_shade():
    res = np.broadcast_to(100.0 + 0.25 * np.arange(DIM, dtype=np.float32), (DIM, DIM)).copy()
    origin = (-1024.0, -512.0)
    sun = (0.6, 0.64, 0.48)
    fr = bytes(_frame(sun, origin=origin))
    sunf = np.array(sun, np.float32).astype(np.float64)
    rng = np.random.default_rng(3)
    worst = 0.0
    for _ in range(400):
        px = float(rng.uniform(-900.0, 800.0))
        pz = float(rng.uniform(-400.0, 900.0))
        d = float(rng.choice([0.0, 220.0, 48.0]))
        qx, qy0 = px + sunf[0] * d, sunf[1] * d
        e = 100.0 + 0.25 * min(max((qx - origin[0]) / TEXEL - 0.5, 0.0), DIM - 1.0)
        py = float(e - qy0 + rng.uniform(-5.0, 5.0))  # q.y within the ramp's reach of e
        want = _smoothstep64(e - SOFT, e + SOFT, py + qy0)
        got = O.terrain_shade(res, fr, np.array([px, py, pz], np.float32), d)
        worst = max(worst, abs(got - want))
    assert worst < 2e-4, worst


def test_oracle_terrain_shade_on_a_ramp_is_its_closed_form():
    _ramp_shade()


# ---- resolve ----

def _pillar(shaft_value=None, blocked=False, legacy=False):
    dense, v, sc, vis, sun, flags, spec_hi = RP._scene(legacy)
    if shaft_value is None:
        return dense, v, sc, vis, sun, flags, spec_hi, RP._resolve(sc, v, flags, spec_hi, vis, sun, RP._atlas())
    v.frame.flags |= DISTANT
    v.frame.shaft_origin[0], v.frame.shaft_origin[1] = -1024.0, -1024.0
    v.frame.shaft_texel, v.frame.shaft_soft = float(TEXEL), SOFT
    field = np.full((DIM, DIM), shaft_value, np.float32)
    if blocked:
        sun = np.full_like(sun, 0x3C00)
    out = O.resolve(sc, v.frame_bytes, O.rspec_from(flags | DISTANT, spec_hi, shaft=field), RP._atlas(), RP.SIZE,
                    RP.LAYERS, RP.MIPS, RP._faces(), vis, sun)
    return dense, v, sc, vis, sun, flags, spec_hi, out


def test_resolve_with_the_field_below_or_above_everything():
    *_, base = _pillar()
    *_, low = _pillar(-1.0e6)
    assert np.array_equal(low.view(np.uint32), base.view(np.uint32))
    *_, high = _pillar(1.0e6)
    dense, v, sc, vis, sun, flags, spec_hi = RP._scene()
    blocked = RP._resolve(sc, v, flags, spec_hi, vis, np.full_like(sun, 0x3C00), RP._atlas())
    assert np.array_equal(high.view(np.uint32), blocked.view(np.uint32))
    assert not np.array_equal(high.view(np.uint32), base.view(np.uint32))


def _floor_misses(legacy=False):
    dense, v, sc, vis, sun, flags, spec_hi = RP._scene(legacy)
    f = v.frame
    qy = 20.0 + float(np.float32(f.sun_dir[1])) * float(np.float32(f.shadow_dist))
    e = qy + 1.3  # the floor's query point sits on the field's edge ramp
    *_, out = _pillar(e, legacy=legacy)
    out = out.reshape(-1, 4)
    s = _smoothstep64(e - SOFT, e + SOFT, qy)
    assert 0.05 < s < 0.95
    lin = [RP._srgb64(c) for c in RP.LEVEL_RGB]
    bad = n = 0
    for i, t, hit, vox, rd in RP._floor_pixels(dense, v, vis):
        if not RP._open(dense, vox[0], vox[2]):
            continue
        lit = 0.0 if float(np.array([sun[i]], np.uint16).view(np.float16)[0]) > 0 else 1.0
        lod = RP._lod(v, t, rd[1])
        l0 = int(math.floor(lod))
        l1 = min(l0 + 1, RP.MIPS - 1)
        albedo = lin[l0] * (1.0 - (lod - l0)) + lin[l1] * (lod - l0)
        want = albedo * RP._light(v, lit * s, 1.0, legacy)
        n += 1
        if np.abs(out[i, :3] - want).max() > 2e-5 * max(1.0, np.abs(want).max()):
            bad += 1
    assert n > 400, n
    return bad


@pytest.mark.parametrize("legacy", [False, True])
def test_floor_pixels_under_the_field_edge_closed_form(legacy):
    assert _floor_misses(legacy) == 0


def _fails(fn, *args):
    try:
        fn(*args)
    except AssertionError:
        return True
    return False


@pytest.mark.parametrize("plant,check", [
    (48, "scan"),    # the scan carries heights from away from the sun
    (49, "scan"),    # the climb without the texel size
    (50, "shade"),   # texel centres at origin + i texel
    (51, "shade"),   # the beyond point on the wrong side
    (52, "floor"),   # the resolve queries at the hit instead of shadow_dist beyond it
])
def test_planted_shaft_mistakes_are_caught(plant, check, tmp_path):
    with O.planted(plant, tmp_path):
        if check == "scan":
            assert _fails(_spike) or _fails(_scan_vs_envelope)
        elif check == "shade":
            assert _fails(_ramp_shade)
        else:
            assert _fails(_floor_misses) or _floor_misses() > 0
