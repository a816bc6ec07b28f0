"""Soft shadows: the oracle's shadow_horizontal / shadow_vertical passes (shadow.wgsl:49-116)
and the resolve's use of their mask, against closed forms of their impulse response.

Input: the floor-and-pillar view with a synthetic primary_shadow word field -- one floor
pixel (x0, y0) occluded at blocker distance D, every other pixel lit. Every pixel within
eight of it is a floor pixel (one plane, one normal, all guides valid), so with
R = min(D SUN_ANGLE SOFT_PENUMBRA proj / t, 4) (t the delta pixel's depth,
proj = H / (2 tan_half_fov); R = 0 below 0.5) and w_R(k) = exp(-2 k^2 / R^2),
S = sum over 0 < |k| <= R of w_R(k):
  row pass: (x0, y0) -> (1 / (1 + S), R); (x0 + q, y0), 0 < |q| <= R -> (w(q) / (1 + w(q)), R);
            every other pixel -> (0, 0);
  column pass: (x0, y0) -> 1 / (1 + S)^2; (x0 + q, y0) -> row(q) / (1 + S);
            (x0 + q, y0 + p), 0 < |p| <= R -> row(q) w(p) / (1 + w(p)); else 0;
  mask = round(255 (1 - column)), to within one level (the f16 row texture).
At R = 4 (a far blocker) and at R = 2.5 the mask matches this; below R = 0.5 it is the
hard mask. On a floor with a one-block step, a delta at the step's foot leaves every
pixel of the upper floor at 255 (the plane test: other planes are not compatible). On
the pillar view's own shadow words the radii stay under one pixel: the mask is the hard
one. The resolve's lane-0 colour is affine in the sun visibility: with the mask on, each
floor pixel's colour is the hard-lit colour mixed toward the hard-occluded one by
1 - mask / 255. Planted mistakes 63 (every neighbour compatible) and 64 (a weight of
exp(-2 |k| / R^2)) fail these.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from tests import test_oracle_resolve_pins as RP
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.world import World, build_view

SOFT = 65536
SUN_ANGLE, SOFT_PENUMBRA = float(np.float32(0.0093)), 1.5
X0, Y0 = 20, 54


def _rspec(flags, spec_hi):
    return O.rspec_from(flags, spec_hi | SOFT)


def _delta_sun(shape, x0, y0, d):
    s = np.zeros(shape, np.float16)
    s[y0, x0] = d
    return s.reshape(-1).view(np.uint16)


def _depth(v, x, y):
    rd = RP._rd64(v.frame)[y, x]
    return (20.0 - float(v.frame.cam_pos[1])) / rd[1]


def _proj(f):
    return float(f.res[1]) / (2.0 * max(float(f.tan_half_fov), 0.001))


def _want_mask(shape, x0, y0, r):
    want = np.full(shape, 255.0)
    if r < 0.5:
        want[y0, x0] = 0.0
        return want
    r = min(r, 4.0)
    w = lambda k: math.exp(-2.0 * k * k / (r * r))
    s = sum(w(k) for k in range(-4, 5) if k != 0 and abs(k) <= r)
    row = {0: 1.0 / (1.0 + s)}
    for q in range(-4, 5):
        if q != 0 and abs(q) <= r:
            row[q] = w(q) / (1.0 + w(q))
    for q, hq in row.items():
        want[y0, x0 + q] = 255.0 * (1.0 - hq / (1.0 + s))  # q = 0: hq / (1 + S) = 1 / (1 + S)^2
        for p in range(-4, 5):
            if p != 0 and abs(p) <= r:
                want[y0 + p, x0 + q] = 255.0 * (1.0 - hq * w(p) / (1.0 + w(p)))
    return want


def _delta_misses(target_r):
    dense, v, sc, vis, sun, flags, spec_hi = RP._scene()
    f = v.frame
    shape = (int(f.res[1]), int(f.res[0]))
    t = _depth(v, X0, Y0)
    scale = SUN_ANGLE * SOFT_PENUMBRA * _proj(f) / t
    d = float(np.float16(min(target_r / scale, 60000.0)))
    r = d * scale if d * scale >= 0.5 else 0.0
    _, mask = O.soft_mask(sc, v.frame_bytes, _rspec(flags, spec_hi), vis, _delta_sun(shape, X0, Y0, d))
    want = _want_mask(shape, X0, Y0, r)
    return int((np.abs(mask.astype(np.float64) - want) > 1.0).sum()), mask, want


@pytest.mark.parametrize("target_r", [1000.0, 2.5, 0.3])
def test_delta_response_closed_form(target_r):
    bad, mask, want = _delta_misses(target_r)
    assert bad == 0, (mask[Y0 - 5:Y0 + 6, X0 - 5:X0 + 6], want[Y0 - 5:Y0 + 6, X0 - 5:X0 + 6].round(1))
    if target_r >= 2.5:
        assert ((mask > 0) & (mask < 255)).sum() >= 20  # a penumbra, not a hard pixel


def test_own_shadow_words_stay_hard():
    dense, v, sc, vis, sun, flags, spec_hi = RP._scene()
    h, mask = O.soft_mask(sc, v.frame_bytes, _rspec(flags, spec_hi), vis, sun)
    occ = sun.view(np.float16).reshape(mask.shape) > 0
    assert occ.sum() > 100
    assert np.array_equal(mask, np.where(occ, 0, 255).astype(np.uint8))
    assert float(h[..., 1].max()) < 1.0


def _step_scene():
    def fn(x, y, z):
        return np.where((y < 20) | ((y < 21) & (z >= 44)), 1, 0)
    world = World()
    world.insert_dense((0, 0, 0), H.dense_from(fn))
    flags = A.FLAG_SHADOWS | 8
    cam = H.camera((32.5, 40.3, 2.5), 90.0, -24.0, RP.W_, RP.H_, flags, sea=10.0)
    v = build_view(world, cam)
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
    vis, _, _, _ = O.render(sc, v.frame_bytes, O.spec_from(flags, A.FLAG_HI_FULL_MARCH),
                            np.zeros(v.tiles[0] * v.tiles[1], np.uint32))
    return v, sc, vis.reshape(RP.H_, RP.W_), flags


def _step_misses():
    v, sc, vis, flags = _step_scene()
    vy = (vis.astype(np.int64) & 0xFFFFFF) >> 8 & 63
    nid = vis.astype(np.int64) >> 37 & 7
    upper = (vis != 0) & (nid == 3) & (vy == 20)
    lower = (vis != 0) & (nid == 3) & (vy == 19)
    x0 = RP.W_ // 2
    rows = [y for y in range(RP.H_ - 1) if upper[y, x0] and not upper[y + 1, x0]]
    assert rows, "no step edge in the view"
    y0 = next(y for y in range(rows[0] + 1, RP.H_) if lower[y, x0])
    assert y0 - rows[0] <= 3
    _, mask = O.soft_mask(sc, v.frame_bytes, _rspec(flags, A.FLAG_HI_FULL_MARCH), vis.reshape(-1),
                          _delta_sun((RP.H_, RP.W_), x0, y0, 1000.0))
    near = np.zeros_like(upper)
    near[y0 - 4:y0 + 5, x0 - 4:x0 + 5] = True
    assert (upper & near).sum() >= 10 and ((mask < 255) & lower & near).sum() >= 20
    return int(((mask != 255) & upper).sum())


def test_other_planes_stay_out_of_the_filter():
    assert _step_misses() == 0


def _affine_misses():
    dense, v, sc, vis, sun, flags, spec_hi = RP._scene()
    f = v.frame
    shape = (int(f.res[1]), int(f.res[0]))
    t = _depth(v, X0, Y0)
    d = float(np.float16(3.0 * t / (SUN_ANGLE * SOFT_PENUMBRA * _proj(f))))
    delta = _delta_sun(shape, X0, Y0, d)
    atlas, faces = RP._atlas(), RP._faces()
    go = lambda spec_hi_, s: O.resolve(sc, v.frame_bytes, O.rspec_from(flags, spec_hi_), atlas, RP.SIZE, RP.LAYERS,
                                      RP.MIPS, faces, vis, s).reshape(-1, 4)[:, :3].astype(np.float64)
    c1 = go(spec_hi, np.zeros_like(delta))
    c0 = go(spec_hi, np.full_like(delta, np.float16(d).view(np.uint16)))
    cs = go(spec_hi | SOFT, delta)
    _, mask = O.soft_mask(sc, v.frame_bytes, _rspec(flags, spec_hi), vis, delta)
    m = mask.reshape(-1).astype(np.float64)[:, None] / 255.0
    want = c0 + (c1 - c0) * m
    floor = np.array([int(k) != 0 and ((int(k) >> 37) & 7) == 3 for k in vis])
    err = np.abs(cs - want).max(axis=1) / np.maximum(1.0, np.abs(want).max(axis=1))
    assert ((m[:, 0] > 0) & (m[:, 0] < 1)).sum() >= 20
    return int((err[floor] > 2e-6).sum())


def test_resolve_mixes_by_the_mask():
    assert _affine_misses() == 0


def _fails(fn, *args):
    try:
        r = fn(*args)
    except AssertionError:
        return True
    return r


@pytest.mark.parametrize("plant,check", [(63, "step"), (64, "delta")])
def test_planted_soft_shadow_mistakes_are_caught(plant, check, tmp_path):
    with O.planted(plant, tmp_path):
        if check == "step":
            assert _fails(_step_misses)
        else:
            assert _fails(lambda: _delta_misses(1000.0)[0])
