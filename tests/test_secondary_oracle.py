"""Pins of the oracle's secondary legs (resolve.wgsl shade_water / shade_glass legs traced
with trace_world, common.wgsl:1205-1253) against an independent float64 construction.

One chunk holds a stone floor, a water layer above it, a submerged stone ledge, a stone
wall, a stone pillar, two low stone islands, a raised block far along the sun direction and
a glass column. For every pixel whose primary hit is the water surface or glass, the expected leg
is built in float64 from the camera and the geometry alone:
  water reflect   mirror of the ray about the +y face normal, from 0.02 above the surface;
  water refract   Snell's law with eta = 1/1.333 (the WGSL refract formula), from 0.02
                  along the refracted ray;
  glass transmit  the primary ray continued from 0.02 past the glass face;
  glass reflect   the mirror about the face normal turned toward the ray;
and its first hit is found with a unit-voxel DDA over the voxels that leg kind cannot skip
(water legs: stone; glass legs: stone and water). Where the transmit leg ends on water
(the glass column stands in the water), shade_water's two legs from that point are built
the same way as the water legs above, from the face the transmit leg hit.

The sun test at each hit (shade_hit, resolve.wgsl:741-745, with shadow_ray,
common.wgsl:1098-1148) is built the same way: from hit + 0.02 n when dot(n, sun) > 0, a
rising ray that starts below the sea level first moves to the sea surface (and is lit if
that is beyond its range), then a unit DDA over stone (the sun ray skips water and glass)
looks for a blocker within the remaining range. Origins whose answer changes under a
1e-4 nudge are left out as ambiguous at f32 precision.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.world import World, build_view

STONE, WATER, GLASS = 1, 13, 10


def _dense():
    def fn(x, y, z):
        out = np.where(y < 20, STONE, np.where(y < 32, WATER, 0))
        out = np.where((x < 24) & (z >= 52) & (z < 56) & (y >= 32) & (y < 52), STONE, out)
        out = np.where((x >= 36) & (x < 50) & (z >= 58) & (y >= 49) & (y < 60), STONE, out)
        out = np.where((x >= 20) & (x < 24) & (z >= 36) & (z < 40) & (y >= 32) & (y < 44), GLASS, out)
        out = np.where((x >= 4) & (x < 8) & (z >= 30) & (z < 40) & (y >= 32) & (y < 48), STONE, out)
        out = np.where((x >= 14) & (x < 28) & (z >= 40) & (z < 52) & (y >= 32) & (y < 35), STONE, out)
        out = np.where((x >= 30) & (x < 40) & (z >= 28) & (z < 38) & (y >= 24) & (y < 28), STONE, out)
        out = np.where((x >= 44) & (x < 54) & (z >= 30) & (z < 44) & (y >= 32) & (y < 40), STONE, out)
        return out
    return H.dense_from(fn)


def _rd64(f, w, h):
    fwd, right, up = (np.array(v, np.float64) for v in (f.cam_fwd, f.cam_right, f.cam_up))
    px, py = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
    nx = px / w * 2.0 - 1.0
    ny = 1.0 - py / h * 2.0
    d = fwd + right * (nx * f.tan_half_fov * f.aspect)[..., None] + up * (ny * f.tan_half_fov)[..., None]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return d.reshape(-1, 3)


def _sun_blocked(stone, o, sun, sea, dist, jump):
    o = np.array(o, np.float64)
    if jump and sun[1] > 1e-4 and o[1] < sea:
        t_surf = (sea - o[1]) / sun[1]
        if t_surf >= dist:
            return False
        o = o + sun * t_surf
        dist = dist - t_surf
    h = H.brute_dda(stone, o, sun)
    return h is not None and h[2] < dist


def _sun_expect(stone, p, nid, sun, sea, dist, jump):
    """None when no sun ray runs, else whether it is blocked; "?" when ambiguous."""
    n = np.zeros(3)
    n[nid >> 1] = -1.0 if nid & 1 == 0 else 1.0  # even ids: the ray moved + along the axis
    if not np.dot(n, sun) > 0.0:
        return None
    o = np.asarray(p, np.float64) + n * 0.02
    answers = {_sun_blocked(stone, o + d, sun, sea, dist, jump)
               for d in [np.zeros(3)] + [e * s for e in np.eye(3) * 1e-4 for s in (1.0, -1.0)]}
    return answers.pop() if len(answers) == 1 else "?"


def _check_sun(leg, exp, tag, tally):
    info = int(leg["info"])
    tested, blocked = bool(info & (1 << 29)), bool(info & (1 << 28))
    if exp == "?":
        tally["ambiguous"] += 1
        return
    if exp is None:
        assert not tested and not blocked, tag
        return
    assert tested, tag
    assert blocked == exp, (tag, blocked, exp)
    tally["blocked" if exp else "lit"] += 1


def _check(leg, exp, tag):
    hit = bool(leg["info"] >> 31)
    assert (leg["info"] >> 30) & 1, tag
    if exp is None:
        assert not hit, tag
        return
    assert hit, (tag, exp)
    v = (int(leg["voxel"]) & 63, (int(leg["voxel"]) >> 6) & 63, (int(leg["voxel"]) >> 12) & 63)
    assert v == exp[0], (tag, v, exp)
    if exp[1] is not None:
        assert (int(leg["info"]) >> 13) & 7 == exp[1], (tag, exp)
        assert abs(float(leg["t"]) - exp[2]) <= 1e-3 * max(1.0, exp[2]), (tag, float(leg["t"]), exp)


@pytest.mark.parametrize("glass_reflect,shadow_cut,legacy_water", [(False, True, False), (True, False, False),
                                                                   (True, False, True)])
def test_secondary_legs_match_float64_construction(glass_reflect, shadow_cut, legacy_water):
    dense = _dense()
    world = World()
    world.insert_dense((0, 0, 0), dense)
    grid3 = dense.reshape(64, 64, 64)  # [y, z, x]
    stone = grid3 == STONE
    stone_or_water = (grid3 == STONE) | (grid3 == WATER)
    w, h = 96, 64
    flags = A.FLAG_WATER_REFLECT | A.FLAG_WATER_REFRACT | A.FLAG_GLASS | A.FLAG_SHADOWS
    if shadow_cut:
        flags |= A.FLAG_WATER_SHADOW_CUT
    sea = 32.0
    cam = H.camera((30.5, 47.3, 6.5), 90.0, -25.0, w, h, flags, sea=sea)
    v = build_view(world, cam)
    spec_hi = A.FLAG_HI_FULL_MARCH | (A.FLAG_HI_GLASS_REFLECT if glass_reflect else 0)
    if legacy_water:  # SPEC_WATER_REPAIR off: no sea-surface jump in shadow_ray
        spec_hi |= A.FLAG_HI_LEGACY_WATER
    sp = O.spec_from(flags, spec_hi)
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
    vis, _, _, _ = O.render(sc, v.frame_bytes, sp, np.zeros(v.tiles[0] * v.tiles[1], np.uint32), shadow=False)
    legs, st = O.secondary(sc, v.frame_bytes, sp, vis)
    f = v.frame
    cam_pos = np.array(f.cam_pos, np.float64)
    rd = _rd64(f, w, h)
    eta = 1.0 / 1.333
    sun = np.array(f.sun_dir, np.float64)
    refract_dist = 16.0 if shadow_cut else 48.0
    counts = {"water": 0, "glass": 0, "glass_water": 0}
    tally = {k: {"lit": 0, "blocked": 0, "ambiguous": 0} for k in ("refract", "reflect", "glass", "behind_glass")}

    def sun_of(leg, ro, rd, exp, dist, kind, tag):
        if exp is None or exp[1] is None:
            return
        if kind == "glass" and int(leg["material"]) == WATER:
            _check_sun(leg, None, tag + " sun", tally[kind])  # shade_water's: no sun ray here
            return
        p = np.asarray(ro, np.float64) + np.asarray(rd, np.float64) * exp[2]
        _check_sun(leg, _sun_expect(stone, p, exp[1], sun, sea, dist, not legacy_water), tag + " sun", tally[kind])

    def water_pair(p, d, nid, rec_refract, rec_reflect, tag, kind_r, kind_m):
        """shade_water's legs from water point p hit along d on face nid (flat normal)."""
        n = np.zeros(3)
        n[nid >> 1] = -1.0 if nid & 1 == 0 else 1.0
        m = d - 2.0 * np.dot(n, d) * n
        e = H.brute_dda(stone, p + n * 0.02, m)
        _check(rec_reflect, e, tag + " reflect")
        sun_of(rec_reflect, p + n * 0.02, m, e, float(f.shadow_dist), kind_m, tag + " reflect")
        dd = np.dot(n, d)
        k = 1.0 - eta * eta * (1.0 - dd * dd)
        r = eta * d - (eta * dd + np.sqrt(k)) * n
        e = H.brute_dda(stone, p + r * 0.02, r)
        _check(rec_refract, e, tag + " refract")
        sun_of(rec_refract, p + r * 0.02, r, e, refract_dist, kind_r, tag + " refract")

    for i in range(w * h):
        key = int(vis[i])
        if key == 0:
            assert legs[i]["info"].sum() == 0
            continue
        low = key & 0xFFFFFF
        vox = ((low >> 16) & 63, (low >> 8) & 63, low & 63)
        nid = (key >> 37) & 7
        mat = dense[vox[0] | (vox[2] << 6) | (vox[1] << 12)]
        d = rd[i]
        tag = f"pixel {i} voxel {vox} face {nid}"
        if mat == WATER:
            assert nid == 3 and vox[1] == 31, tag  # the surface, seen from above (face -y entry)
            counts["water"] += 1
            t = (32.0 - cam_pos[1]) / d[1]
            p = cam_pos + d * t
            water_pair(p, d, 3, legs[i][0], legs[i][1], tag, "refract", "reflect")
            assert legs[i][2]["info"] == 0 and legs[i][3]["info"] == 0, tag
        elif mat == GLASS:
            counts["glass"] += 1
            axis = nid >> 1
            plane = vox[axis] + (0 if nid & 1 == 0 else 1)
            t = (plane - cam_pos[axis]) / d[axis]
            p = cam_pos + d * t
            e = H.brute_dda(stone_or_water, p + d * 0.02, d)
            _check(legs[i][0], e, tag + " transmit")
            sun_of(legs[i][0], p + d * 0.02, d, e, 48.0, "glass", tag + " transmit")
            ev = None if e is None else e[0]
            if ev is not None and dense[ev[0] | (ev[2] << 6) | (ev[1] << 12)] == WATER:
                assert e[1] is not None, tag
                counts["glass_water"] += 1
                p2 = p + d * 0.02 + d * e[2]
                water_pair(p2, d, e[1], legs[i][2], legs[i][3], tag + " behind glass", "behind_glass",
                           "behind_glass")
            else:
                assert legs[i][2]["info"] == 0 and legs[i][3]["info"] == 0, tag
            if glass_reflect:
                n = np.zeros(3)
                n[axis] = -1.0 if d[axis] > 0 else 1.0
                m = d - 2.0 * np.dot(n, d) * n
                e = H.brute_dda(stone_or_water, p + n * 0.02, m)
                _check(legs[i][1], e, tag + " glass reflect")
                sun_of(legs[i][1], p + n * 0.02, m, e, 48.0, "glass", tag + " glass reflect")
            else:
                assert legs[i][1]["info"] == 0, tag
        else:
            assert legs[i]["info"].sum() == 0, tag
    assert counts["water"] > 200 and counts["glass"] > 10 and counts["glass_water"] >= 5, counts
    assert st["rays_secondary"] == (2 * counts["water"] + counts["glass"] * (2 if glass_reflect else 1)
                                    + 2 * counts["glass_water"])
    print(tally)
    # both outcomes occur for every leg kind
    print(counts)
    for kind in ("refract", "reflect", "glass"):
        assert tally[kind]["lit"] >= 5 and tally[kind]["blocked"] >= 5, tally
    assert sum(t["ambiguous"] for t in tally.values()) < 10, tally
