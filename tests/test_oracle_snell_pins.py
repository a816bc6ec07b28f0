"""The oracle's Snell bend (march.wgsl:23-45) against a float64 construction.

One chunk: a stone floor, water up to the sea level at y = 32, and a stone slab in the air
above it (y in [40, 42), z >= 27). An underwater camera looks up through the surface.
For every pixel the expected first hit is built in float64 from the camera alone: the
straight ray's first stone voxel (water is skipped under water); if it misses and goes
up, the ray is refracted at the surface with the WGSL refract formula (eta = 1.333 about
(0, -1, 0); none past the critical angle) and a unit-voxel DDA from 0.02 past the surface
finds the slab, at distance t_surf + t2. The key's voxel and face must be that hit's and
its depth field depth_key(t) within 3 units. Rays whose answer changes under a 1e-4 nudge
are left out. Planted mistakes (VXO_PLANT 16: eta inverted, 17: t without t_surf) fail.
"""
import functools
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.world import World, build_view

STONE, WATER = 1, 13
W_, H_ = 64, 48
SEA = 32.0
SNELL = 256


def _world():
    def fn(x, y, z):
        out = np.where(y < 8, STONE, np.where(y < 32, WATER, 0))
        return np.where((y >= 40) & (y < 42) & (z >= 27), STONE, out)
    dense = H.dense_from(fn)
    w = World()
    w.insert_dense((0, 0, 0), dense)
    return w, dense.reshape(64, 64, 64)


def _refract(i, n, eta):
    d = float(np.dot(n, i))
    k = 1.0 - eta * eta * (1.0 - d * d)
    if k < 0.0:
        return None
    return eta * i - (eta * d + math.sqrt(k)) * n


def _expect(stone, cam, rd, far):
    h = H.brute_dda(stone, cam, rd)
    if h is not None:
        return h[0], h[1], h[2], False
    if rd[1] <= 0.0:
        return None
    t_surf = (SEA - cam[1]) / rd[1]
    if t_surf >= far:
        return None
    rd2 = _refract(rd, np.array([0.0, -1.0, 0.0]), 1.333)
    if rd2 is None:
        return None
    surf = cam + rd * t_surf
    h2 = H.brute_dda(stone, surf + rd2 * 0.02, rd2)
    if h2 is None:
        return None
    return h2[0], h2[1], t_surf + h2[2], True


def _setup():
    world, grid = _world()
    flags = A.FLAG_UNDERWATER | A.FLAG_HIZ
    cam = H.camera((32.5, 20.3, 6.5), 90.0, 50.0, W_, H_, flags, sea=SEA)
    return world, grid, flags, build_view(world, cam)


@functools.lru_cache(maxsize=1)
def _expected():
    """Per pixel: None when ambiguous, else (the float64 hit or None,)."""
    _, grid, _, v = _setup()
    stone = grid == STONE
    f = v.frame
    far = float(f.far_)
    camp = np.array(f.cam_pos, np.float64)
    fwd, right, up = (np.array(x, np.float64) for x in (f.cam_fwd, f.cam_right, f.cam_up))
    out = []
    for py in range(H_):
        for px in range(W_):
            nx = (px + 0.5) / W_ * 2.0 - 1.0
            ny = 1.0 - (py + 0.5) / H_ * 2.0
            d = fwd + right * (nx * f.tan_half_fov * f.aspect) + up * (ny * f.tan_half_fov)
            rd = d / np.linalg.norm(d)
            e = _expect(stone, camp, rd, far)
            alts = [_expect(stone, camp + dd, rd, far) for dd in np.eye(3) * 1e-4]
            if any((a is None) != (e is None) or (a and e and (a[0], a[1]) != (e[0], e[1])) for a in alts):
                out.append(None)  # ambiguous at f32 precision
            else:
                out.append((e,))
    return out, far


def _misses():
    world, _, flags, v = _setup()
    spec_hi = A.FLAG_HI_FULL_MARCH | SNELL
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
    vis, _, _, _ = O.render(sc, v.frame_bytes, O.spec_from(flags, spec_hi), np.zeros(v.tiles[0] * v.tiles[1], np.uint32),
                            shadow=False)
    exp, far = _expected()
    bad = bent = n = 0
    for i, ex in enumerate(exp):
        if ex is None:
            continue
        e = ex[0]
        key = int(vis[i])
        n += 1
        if e is None:
            bad += key != 0
            continue
        bent += e[3]
        if key == 0:
            bad += 1
            continue
        low = key & 0xFFFFFF
        vox = ((low >> 16) & 63, (low >> 8) & 63, low & 63)
        depth = key >> 40
        want = int(min(max(1.0 - e[2] / far, 0.0), 1.0) * 16777215.0)
        if vox != tuple(e[0]) or ((key >> 37) & 7) != e[1] or abs(depth - want) > 3:
            bad += 1
    assert n > 2000 and bent > 100, (n, bent)
    return bad


def test_snell_bend_matches_float64_construction():
    assert _misses() == 0


@pytest.mark.parametrize("plant", [16, 17])
def test_planted_snell_mistakes_are_caught(plant, tmp_path):
    with O.planted(plant, tmp_path):
        assert _misses() > 0
