"""Ambient probes: the host bake (vxw_probe_*, src/probe.rs), the probe field's layout
(write_probe, src/render/mod.rs:1398-1445) and the oracle's probe_field and cube / bounce /
sun terms in the resolve (resolve.wgsl:95-118, :552-576).

Host bake, ports of tests/probe.rs on this build's generator (its heights and surface
blocks replace WorldGen's): an open probe's cube is 1 and a buried one falls back to 1, a
probe in the ramp is lifted off its raw occlusion without reaching 1, a one-sided horizon
darkens only the face that looks at it, grass ground leaves the floor's luminance alone
while snow and water carry theirs, over mixed ground the bounce varies across a chunk and
between axes with the brighter ground on the brighter axis, shadowed ground bounces no
more than open ground and leaves the occlusion alone, and chunks in one upload window
never share a block of the torus.
Oracle probe_field, on a field whose channels are linear in the texel indices (exact in
f16): trilinear filtering returns the linear function at texel coordinates p / 4 - 0.5
(the slab of the normal, y kept half a texel inside it), wraps in x and z, clamps in y.
Resolve, on the floor-and-pillar chunk: a field of ones with the sun bounce off changes
no bit; a constant field (rgb B, alpha a) gives the lane-0 closed form with the face shade
times a and the floor light FLOOR (ambient B + B 0.10) under both lighting forms.
The probe tap (resolve.wgsl:120-127) on the same linear field returns the linear function
at the clamped texel coordinates of fract(p / 512) over the whole texture (clamp to edge on
every axis); cross faces sample the +y slab. Under SPEC_PROBE_TAP with a constant field B
the floor's sky ambient is scaled by B; under SPEC_PROBE_AMBIENT the ambient is B itself,
the face shade 1 and the floor light 0.
Planted mistakes 57-61, 65 and 66 each fail one of these.
"""
import dataclasses
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import test_oracle_resolve_pins as RP
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.world import (albedo, probe_bake, probe_bounce_of_uniform_ground, probe_cube_from_tangents,
                                       probe_uploadable, surface_block)

CUBE, BOUNCE, SUN, TAP, AMBIENT = 67108864, 134217728, 536870912, 16777216, 33554432
GRASS, SNOW, WATER = 3, 14, 13
FLOOR = np.array([0.85, 0.90, 1.00])


def _luma(c):
    return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]


def _gen():
    return dataclasses.replace(W.WORKLOADS[1].gen, seed=1337)


# ---- host bake (ports of tests/probe.rs) ----

def test_open_buried_and_ramp_probes():
    assert np.allclose(probe_cube_from_tangents(np.zeros(16)), 1.0, atol=1e-4)
    assert np.allclose(probe_cube_from_tangents(np.full(16, 64.0)), 1.0, atol=1e-3)
    slot = probe_cube_from_tangents(np.full(16, 3.0))
    assert 0.2 < slot[3] < 0.95


def test_a_one_sided_horizon_darkens_only_the_face_that_looks_at_it():
    tan = np.zeros(16)
    tan[[0, 1, 2, 14, 15]] = 4.0  # a wall in the +X azimuths
    c = probe_cube_from_tangents(tan)
    assert c[1] < c[0] - 0.05 and c[3] < 1.0 and abs(c[0] - 1.0) < 0.02 and c[2] == 1.0


def test_reference_ground_leaves_the_exposure_alone():
    want = _luma(FLOOR)
    m = probe_bounce_of_uniform_ground(albedo(GRASS))
    for v in m:
        assert abs(_luma(FLOOR * v) - want) < 1e-4
    assert probe_bounce_of_uniform_ground(albedo(SNOW))[3][1] > 2.0
    w = probe_bounce_of_uniform_ground(albedo(WATER))
    assert w[3][2] > 3.0 * w[3][0]


def _mixed_surface_chunk(gen):
    for cz in range(-4, 4):
        for cx in range(-4, 4):
            o = (cx * 64, 128, cz * 64)
            seen = {surface_block(gen, o[0] + i, o[2] + (i * 7) % 64, gen.height(o[0] + i, o[2] + (i * 7) % 64))
                    for i in range(64)}
            if len(seen) >= 2:
                return o
    raise AssertionError("no chunk with two surface block types in the searched range")


@pytest.fixture(scope="module")
def mixed():
    gen = _gen()
    o = _mixed_surface_chunk(gen)
    return gen, o, probe_bake(gen, o, True), probe_bake(gen, o, False)


def test_the_bounce_varies_across_a_chunk_and_between_axes(mixed):
    gen, o, (occl, bounce), _ = mixed
    assert bounce.max() / bounce.min() > 1.5
    lum = _luma(np.moveaxis(bounce, -1, 0))  # (6, pz, py, px)
    horiz = lum[[0, 1, 4, 5]]
    assert (horiz.max(0) - horiz.min(0)).max() > 0.05


def test_the_brighter_ground_is_on_the_brighter_axis(mixed):
    gen, o, (occl, bounce), _ = mixed
    lum = _luma(np.moveaxis(bounce, -1, 0))
    gap = lum[1] - lum[0]
    pz, py, px = np.unravel_index(np.argmax(np.abs(gap)), gap.shape)
    best = gap[pz, py, px]
    assert abs(best) > 0.05
    cx, cz = o[0] + 2 + 4 * px, o[2] + 2 + 4 * pz
    side = []
    for sign in (-1, 1):
        vals = []
        for dx in range(1, 25):
            for dz in range(-24, 25):
                x, z = cx + sign * dx, cz + dz
                vals.append(_luma(albedo(surface_block(gen, x, z, gen.height(x, z)))))
        side.append(np.mean(vals))
    assert (best > 0.0) == (side[1] > side[0]), (best, side)


def test_shadowed_ground_bounces_no_more_than_open_ground(mixed):
    gen, o, (occl_s, shaded), (occl_o, open_) = mixed
    assert np.all(shaded <= open_ + 1e-6)
    assert (np.abs(shaded - open_) > 1e-6).any()
    assert np.array_equal(occl_s, occl_o)  # the second bounce leaves the occlusion cube alone


def test_two_uploadable_chunks_never_share_a_block():
    cam = (1000.5, 80.0, -700.5)
    gx, gz = int(cam[0]) // 64 * 64, -((-int(cam[2])) // 64) * 64
    window = [(gx + ox, 64, gz + oz) for ox in range(-320, 385, 64) for oz in range(-320, 385, 64)
              if probe_uploadable((gx + ox, 64, gz + oz), cam)]
    assert window
    blocks = [(((o[0] // 4) % 128), o[1] // 4, ((o[2] // 4) % 128)) for o in window]
    assert len(set(blocks)) == len(blocks)
    near, far = (64, 64, 128), (64 + 512, 64, 128)
    assert ((near[0] // 4) % 128, (near[2] // 4) % 128) == ((far[0] // 4) % 128, (far[2] // 4) % 128)
    assert not probe_uploadable(far, (near[0] + 32.0, 0.0, near[2] + 32.0))


# ---- oracle probe_field ----

def _linear_field():
    z, y, x = np.meshgrid(np.arange(128), np.arange(768), np.arange(128), indexing="ij")
    f = np.stack([x / 8.0, y / 8.0, z / 8.0, (x + z) / 16.0], axis=-1).astype(np.float16)
    return f.view(np.uint16)


def _field_sample_misses():
    field = _linear_field()
    rng = np.random.default_rng(4)
    worst = 0.0
    for _ in range(300):
        nid = int(rng.integers(0, 6))
        p = np.array([rng.uniform(8.0, 500.0), rng.uniform(-10.0, 530.0), rng.uniform(8.0, 500.0)], np.float32)
        got = O.probe_field(field, p, nid)
        tx, tz = float(p[0]) / 4.0 - 0.5, float(p[2]) / 4.0 - 0.5
        y = min(max(float(p[1]) / 4.0, 0.5), 127.5)
        ty = nid * 128 + y - 0.5
        want = np.array([tx / 8.0, ty / 8.0, tz / 8.0, (tx + tz) / 16.0])
        worst = max(worst, float(np.abs(got - want).max()))
    # the seam in x: between texel 127 and texel 0 (512 blocks on)
    p = np.array([512.0 + 1.0, 100.0, 200.0], np.float32)  # tx = 128.25 - 0.5: texel 127 .. 0 at 0.75
    got = O.probe_field(field, p, 2)
    want_r = (127 / 8.0) * 0.25 + 0.0 * 0.75
    worst = max(worst, abs(float(got[0]) - want_r))
    return worst


def test_oracle_probe_field_on_a_linear_field():
    assert _field_sample_misses() < 1e-4


def _tap_misses():
    field = _linear_field()
    rng = np.random.default_rng(6)
    worst = 0.0
    for _ in range(300):
        p = rng.uniform(-1500.0, 1500.0, 3).astype(np.float32)
        got = O.probe_tap(field, p)
        u = np.asarray(p, np.float64) / 512.0
        u = u - np.floor(u)
        t = np.clip(u * np.array([128.0, 768.0, 128.0]) - 0.5, 0.0, [127.0, 767.0, 127.0])
        worst = max(worst, float(np.abs(got - t / 8.0).max()))
    for nid in (6, 7):  # cross faces sample the +y slab
        p = np.array([100.3, 60.7, 222.9], np.float32)
        worst = max(worst, float(np.abs(O.probe_field(field, p, nid) - O.probe_field(field, p, 3)).max()))
    return worst


def test_oracle_probe_tap_on_a_linear_field():
    assert _tap_misses() < 1e-3


# ---- resolve ----

def _pillar_probe(field, spec_extra, legacy=False):
    dense, v, sc, vis, sun, flags, spec_hi = RP._scene(legacy)
    f = flags | spec_extra
    out = O.resolve(sc, v.frame_bytes, O.rspec_from(f, spec_hi, probe=field), RP._atlas(), RP.SIZE, RP.LAYERS,
                    RP.MIPS, RP._faces(), vis, sun)
    return dense, v, vis, sun, out


def _constant_field(rgb, a):
    f = np.empty((128, 768, 128, 4), np.float16)
    f[..., :3] = rgb
    f[..., 3] = a
    return f.view(np.uint16)


def test_a_field_of_ones_changes_nothing():
    dense, v, sc, vis, sun, flags, spec_hi = RP._scene()
    base = RP._resolve(sc, v, flags, spec_hi, vis, sun, RP._atlas())
    *_, ones = _pillar_probe(_constant_field(1.0, 1.0), CUBE | BOUNCE)
    assert np.array_equal(ones.view(np.uint32), base.view(np.uint32))


def _floor_probe_misses(legacy=False):
    B, a = 0.75, 0.5
    dense, v, vis, sun, out = _pillar_probe(_constant_field(B, a), CUBE | BOUNCE | SUN, legacy)
    out = out.reshape(-1, 4)
    f = v.frame
    d, sy = float(f.daylight), float(f.sun_dir[1])
    lin = [RP._srgb64(c) for c in RP.LEVEL_RGB]
    bad = n = 0
    for i, t, hit, vox, rd in RP._floor_pixels(dense, v, vis):
        if not RP._open(dense, vox[0], vox[2]):
            continue
        lit = 0.0 if float(np.array([sun[i]], np.uint16).view(np.float16)[0]) > 0 else 1.0
        direct = lit * sy * d if sy > 0.0 and d > 0.01 else 0.0
        floor = FLOOR * float(f.ambient) * B + FLOOR * B * 0.10  # sky_sm = 1 on the open floor
        if legacy:
            light = ((0.5 * RP.SKY_TINT * d + RP.SUN_RGB * direct * 0.52) * 1.0 + floor) * (1.0 * a)
        else:
            light = 0.5 * RP.SKY_TINT * 0.6 * d * 1.0 * a + RP.SUN_RGB * direct * 0.78 + floor
        lod = RP._lod(v, t, rd[1])
        l0 = int(math.floor(lod))
        l1 = min(l0 + 1, RP.MIPS - 1)
        alb = lin[l0] * (1.0 - (lod - l0)) + lin[l1] * (lod - l0)
        want = alb * light
        n += 1
        if np.abs(out[i, :3] - want).max() > 2e-5 * max(1.0, np.abs(want).max()):
            bad += 1
    assert n > 400, n
    return bad


@pytest.mark.parametrize("legacy", [False, True])
def test_floor_pixels_under_a_constant_field_closed_form(legacy):
    assert _floor_probe_misses(legacy) == 0


def _floor_tap_misses(mode, legacy=False):
    B = float(np.float16(0.6))  # the field holds f16
    dense, v, vis, sun, out = _pillar_probe(_constant_field(B, 0.5), TAP if mode == "tap" else AMBIENT, legacy)
    out = out.reshape(-1, 4)
    f = v.frame
    d, sy = float(f.daylight), float(f.sun_dir[1])
    lin = [RP._srgb64(c) for c in RP.LEVEL_RGB]
    bad = n = 0
    for i, t, hit, vox, rd in RP._floor_pixels(dense, v, vis):
        if not RP._open(dense, vox[0], vox[2]):
            continue
        lit = 0.0 if float(np.array([sun[i]], np.uint16).view(np.float16)[0]) > 0 else 1.0
        direct = lit * sy * d if sy > 0.0 and d > 0.01 else 0.0
        sky = RP.SKY_TINT * d * (1.0 if legacy else 0.6)  # the gathered ambient on the open floor
        amb = np.full(3, B) if mode == "ambient" else sky * B
        floor = np.zeros(3) if mode == "ambient" else FLOOR * float(f.ambient)
        if legacy:
            light = (0.5 * amb + RP.SUN_RGB * direct * 0.52) * 1.0 + floor
        else:
            light = 0.5 * amb * 1.0 + RP.SUN_RGB * direct * 0.78 + floor
        lod = RP._lod(v, t, rd[1])
        l0 = int(math.floor(lod))
        l1 = min(l0 + 1, RP.MIPS - 1)
        alb = lin[l0] * (1.0 - (lod - l0)) + lin[l1] * (lod - l0)
        want = alb * light
        n += 1
        if np.abs(out[i, :3] - want).max() > 2e-5 * max(1.0, np.abs(want).max()):
            bad += 1
    assert n > 400, n
    return bad


@pytest.mark.parametrize("mode", ["tap", "ambient"])
@pytest.mark.parametrize("legacy", [False, True])
def test_floor_pixels_under_the_probe_tap_closed_form(mode, legacy):
    assert _floor_tap_misses(mode, legacy) == 0


def _fails(fn, *args):
    try:
        r = fn(*args)
    except AssertionError:
        return True
    return r


@pytest.mark.parametrize("plant,check", [
    (57, "field"),   # the slab of the neighbouring normal
    (58, "field"),   # texel coordinates without the half-texel offset
    (59, "field"),   # y not kept inside the slab
    (60, "floor"),   # the bounce left off the floor light
    (61, "floor"),   # the cube left off the face shade
    (65, "tap"),     # the tap's x texel coordinate without the half-texel offset
    (66, "ambient"), # the floor light kept under SPEC_PROBE_AMBIENT
])
def test_planted_probe_mistakes_are_caught(plant, check, tmp_path):
    with O.planted(plant, tmp_path):
        if check == "field":
            assert _fails(lambda: _field_sample_misses() >= 1e-4)
        elif check == "tap":
            assert _fails(lambda: _tap_misses() >= 1e-3)
        elif check == "ambient":
            assert _fails(_floor_tap_misses, "ambient")
        else:
            assert _fails(_floor_probe_misses)
