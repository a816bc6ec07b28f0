"""Pins of the CPU oracle: the reference's own fixtures and closed forms.

Each test fixes a part of the restated WGSL against something independent of the
oracle's code: a reference test's literal values, a published inverse, numpy's IEEE
conversion, a float64 brute-force DDA, or an invariant the reference relies on.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.world import World, build_view, pairs_for

L = O.lib()


def _fb(frame):
    return np.frombuffer(bytes(frame), np.uint8).copy()


# ---- key packing: tests/cutout.rs:16-59 -------------------------------------------------
@pytest.mark.parametrize("v", [[0, 0, 0], [63, 63, 63], [1, 62, 33], [63, 0, 7]])
@pytest.mark.parametrize("m", [[0, 0, 0], [3, 3, 3], [1, 2, 3], [0, 3, 1]])
def test_micro_round_trips_without_disturbing_the_voxel(v, m):
    f = _fb(H.frame_for())
    va, ma = np.array(v, np.uint32), np.array(m, np.uint32)
    key = L.vxo_make_key(C.c_void_p(f.ctypes.data), 10.0, 5, 4321, C.c_void_p(va.ctypes.data),
                         C.c_void_p(ma.ctypes.data))
    low = key & 0xFFFFFF
    b = [(low >> 16) & 0xFF, (low >> 8) & 0xFF, low & 0xFF]
    assert [x & 63 for x in b] == v
    assert [x >> 6 for x in b] == m
    assert (key >> 37) & 7 == 5 and (key >> 24) & 0x1FFF == 4321
    assert key >> 40 == L.vxo_depth_key(C.c_void_p(f.ctypes.data), 10.0)
    if m == [0, 0, 0]:
        assert low == (v[0] << 16) | (v[1] << 8) | v[2]


# ---- depth_key (common.wgsl:1320-1322) ----------------------------------------------------
def test_depth_key_endpoints_and_monotone():
    f = _fb(H.frame_for(far=2600.0))
    fp = C.c_void_p(f.ctypes.data)
    assert L.vxo_depth_key(fp, 0.0) == 16777215
    assert L.vxo_depth_key(fp, -5.0) == 16777215
    assert L.vxo_depth_key(fp, 2600.0) == 0
    assert L.vxo_depth_key(fp, 1e6) == 0
    ts = np.linspace(0, 2600, 5001, dtype=np.float32)
    keys = np.array([L.vxo_depth_key(fp, float(t)) for t in ts])
    assert np.all(np.diff(keys.astype(np.int64)) <= 0)
    # one key step is far / 2^24 world units: 100 units is ~645 k steps
    assert abs(int(L.vxo_depth_key(fp, 0.0)) - int(L.vxo_depth_key(fp, 100.0)) - 645277) <= 1


# ---- pair word / capacity: tests/pairs.rs:13-77 ------------------------------------------
def test_pair_capacity_matches_reference_budget():
    tiles = lambda w, h: ((w + 7) // 8, (h + 7) // 8)
    for w, h in [(1280, 720), (1920, 1080), (2560, 1440)]:
        assert pairs_for(*tiles(w, h)) == A.VX_MAX_PAIRS
    for w, h, demand in [(3840, 2160, 1_313_456), (5120, 2880, 2_325_513)]:
        assert pairs_for(*tiles(w, h)) > demand
    t8k = tiles(7680, 4320)
    assert t8k[0] * t8k[1] <= A.VX_MAX_TILES
    t = tiles(8192, 4320)
    assert t[0] * t[1] > A.VX_MAX_TILES
    assert A.VX_MAX_CHUNKS == (1 << 13) - 1 and A.VX_MAX_TILES == 1 << (32 - 13)


# ---- R16F store ---------------------------------------------------------------------------
def test_f16_conversion_is_ieee_round_to_nearest_even():
    rng = np.random.default_rng(5)
    xs = np.concatenate([
        rng.uniform(0, 300, 20000).astype(np.float32),
        np.float32(2.0) ** rng.uniform(-30, 17, 20000).astype(np.float32),
        np.array([0.001, 65504.0, 65519.99, 65520.0, 6.1035156e-05, 5.96e-08, 2.98e-08, 0.0, 1.0, 220.0],
                 np.float32),
    ]).astype(np.float32)
    got = np.array([L.vxo_f32_to_f16(float(x)) for x in xs], np.uint16)
    want = xs.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)


# ---- hash_u32 is lowbias32 (common.wgsl:1422-1430): its published inverse undoes it ------
def test_hash_u32_inverts_with_published_inverse():
    def inv(x):
        x ^= x >> 16
        x = (x * 0x43021123) & 0xFFFFFFFF
        x ^= (x >> 15) ^ (x >> 30)
        x = (x * 0x1D69E2A5) & 0xFFFFFFFF
        x ^= x >> 16
        return x
    rng = np.random.default_rng(1)
    for x in list(rng.integers(0, 2 ** 32, 2000)) + [0, 1, 0xFFFFFFFF]:
        assert inv(L.vxo_hash_u32(int(x))) == int(x)
    assert L.vxo_hash_u32(0) == 0


# ---- ray_dir inverts view_proj: tests/camera.rs:46-75 (src/math.rs:42-67) -----------------
def _view_proj(pos, fwd, right, up, tan_half, aspect, near, far):
    view = np.array([
        [right[0], right[1], right[2], -np.dot(right, pos)],
        [up[0], up[1], up[2], -np.dot(up, pos)],
        [-fwd[0], -fwd[1], -fwd[2], np.dot(fwd, pos)],
        [0, 0, 0, 1]], np.float64)
    f = 1.0 / tan_half
    proj = np.array([[f / aspect, 0, 0, 0], [0, f, 0, 0],
                     [0, 0, far / (near - far), near * far / (near - far)], [0, 0, -1, 0]], np.float64)
    return proj @ view


@pytest.mark.parametrize("yaw,pitch", [(40.0, -14.0), (0.0, 0.0), (-155.0, 62.0), (91.0, -80.0)])
def test_ray_dir_inverts_view_proj(yaw, pitch):
    from voxelcraft_menu_amd.world import camera_basis
    res = (1280, 720)
    fwd, right, up = camera_basis(math.radians(yaw), math.radians(pitch))
    fr = H.frame_for(*res)
    pos = (31.5, 92.25, -404.0)
    for i in range(3):
        fr.cam_pos[i], fr.cam_fwd[i], fr.cam_right[i], fr.cam_up[i] = pos[i], fwd[i], right[i], up[i]
    fr.tan_half_fov = math.tan(math.radians(70.0) * 0.5)
    fr.aspect = res[0] / res[1]
    m = _view_proj(np.array(pos), np.array(fwd), np.array(right), np.array(up), fr.tan_half_fov, fr.aspect,
                   0.05, 2600.0)
    fb = _fb(fr)
    for px, py in [(0.5, 0.5), (640.5, 360.5), (1279.5, 719.5), (17.5, 700.5)]:
        rd = (C.c_float * 3)()
        L.vxo_ray_dir(C.c_void_p(fb.ctypes.data), px, py, rd)
        rd = np.array(list(rd), np.float64)
        assert abs(np.linalg.norm(rd) - 1.0) < 1e-6
        for t in (0.2, 7.0, 1200.0):
            clip = m @ np.append(np.array(pos) + rd * t, 1.0)
            assert clip[3] > 0
            x = (clip[0] / clip[3] * 0.5 + 0.5) * res[0]
            y = (0.5 - clip[1] / clip[3] * 0.5) * res[1]
            assert abs(x - px) < 0.01 and abs(y - py) < 0.01


# ---- march_chunk vs a float64 unit-step DDA on opaque worlds ------------------------------
def _opaque_world(seed, density):
    rng = np.random.default_rng(seed)
    solid = rng.random((64, 64, 64)) < density  # [y, z, x]
    ids = rng.integers(1, 7, (64, 64, 64)).astype(np.uint16)  # stone..log: all opaque, no cutout
    ids[ids == 6] = 1
    dense = np.where(solid, ids, 0).astype(np.uint16).reshape(-1)
    w = World()
    w.insert_dense((0, 0, 0), dense)
    return w, solid, dense


@pytest.mark.parametrize("seed,density", [(1, 0.002), (2, 0.01), (3, 0.05), (4, 0.2)])
def test_march_first_hit_matches_bruteforce_dda(seed, density):
    w, solid, _ = _opaque_world(seed, density)
    ch = H.gpu_chunk(w, (0, 0, 0))
    sc = H.scene_for(w, ch)
    fr = H.frame_for()
    sp = O.spec_from(A.FLAG_LEAF_CUTOUT | A.FLAG_GLASS, A.FLAG_HI_FULL_MARCH)
    rng = np.random.default_rng(100 + seed)
    checked = hits = 0
    for _ in range(400):
        if rng.random() < 0.5:
            ro = rng.uniform(0.0, 64.0, 3).astype(np.float32)
        else:
            ro = (rng.uniform(-40, 104, 3)).astype(np.float32)
        d = rng.normal(size=3)
        rd = (d / np.linalg.norm(d)).astype(np.float32)
        rd = np.where(np.abs(rd) < 1e-6, np.float32(1e-6), rd).astype(np.float32)
        h = H.march_chunk(sc, fr, sp, 0, ro, rd)
        ref = H.brute_dda(solid, ro.astype(np.float64), rd.astype(np.float64))
        checked += 1
        if ref is None:
            assert not h.hit, (ro, rd)
            continue
        hits += 1
        assert h.hit, (ro, rd, ref)
        assert tuple(h.voxel) == ref[0], (ro, rd, tuple(h.voxel), ref)
        if ref[1] is not None:
            assert h.normal == ref[1]
            assert abs(h.t - ref[2]) <= 1e-4 * max(1.0, ref[2])
    assert hits > (5 if density < 0.005 else 20)


# ---- tree round trip through the oracle's WGSL decode (block_at) --------------------------
def test_oracle_block_at_decodes_host_pools_to_the_dense_grid():
    rng = np.random.default_rng(7)
    w = World()
    denses = []
    for k in range(4):
        # layered terrain + noise, 1..20 distinct ids, water, emitter and cutout ids
        d = H.dense_from(lambda x, y, z: np.where(y < 20 + (x * 7 + z * 3 + k) % 13, 1 + (x + y * 3 + z + k) % 20, 0))
        holes = rng.random(d.size) < 0.3
        d[holes] = 0
        sprinkle = rng.random(d.size) < 0.05
        d[sprinkle] = rng.choice([7, 10, 11, 13, 16], sprinkle.sum())
        denses.append(d)
        w.insert_dense((k, 0, 0), d)
    chunks = np.concatenate([H.gpu_chunk(w, (k, 0, 0)) for k in range(4)])
    sc = H.scene_for(w, chunks)
    for k, d in enumerate(denses):
        idx = rng.integers(0, d.size, 4000)
        for i in idx:
            x, z, y = i & 63, (i >> 6) & 63, i >> 12
            if d[i] == 0:
                continue
            v = np.array([x, y, z], np.uint32)
            got = L.vxo_block_at(C.byref(sc.pools), C.c_void_p(sc.chunks.ctypes.data), k, C.c_void_p(v.ctypes.data))
            assert got == d[i]


# ---- reference tree tests (tests/tree.rs, tests/full_march.rs) ----------------------------
def _terrain(seed):
    def f(x, y, z):
        h = 20 + ((x * 31 + z * 17 + seed) % 9)
        return np.where(y < h - 3, 1, np.where(y < h, 2, np.where(y == h, 3, 0)))
    return H.dense_from(f)


def test_world_roundtrip_and_aabb():
    w = World()
    d = _terrain(1)
    w.insert_dense((3, 1, -2), d)
    assert np.array_equal(w.extract_dense((3, 1, -2)), d)
    info = w.chunk_info((3, 1, -2))
    assert info.aabb_min[1] == 0 and info.aabb_max[1] <= 45
    assert info.aabb_min[0] == 0 and info.aabb_max[0] == 64
    assert w.get_block(0, -1, 0) == 0 and w.get_block(1000, 10, 1000) == 0
    assert w.get_block(3 * 64 + 5, 64 + 2, -2 * 64 + 9) == d[H.dense_index(5, 2, 9)]
    w.remove_chunk((3, 1, -2))
    for p in (0, 1):
        s = w.pool_stats(p)
        assert s["live_elems"] == 0 and s["referenced_elems"] == 0
    assert w.pool_stats(2)["live_elems"] == 0


def test_identical_chunks_share_runs():
    w = World()
    d = _terrain(2)
    w.insert_dense((0, 0, 0), d)
    live = [w.pool_stats(p)["live_elems"] for p in (0, 1)]
    w.insert_dense((5, 0, 5), d)
    assert [w.pool_stats(p)["live_elems"] for p in (0, 1)] == live
    s = w.pool_stats(0)
    assert s["referenced_elems"] >= 2 * s["live_elems"]
    assert np.array_equal(w.extract_dense((5, 0, 5)), d)


def test_uniform_chunk_costs_nothing():
    w = World()
    w.insert_dense((0, 0, 0), np.full(H.VOL, 1, np.uint16))
    info = w.chunk_info((0, 0, 0))
    assert info.root[2] & 0x80000000 and info.attr_uniform and info.attr_base == 1
    assert all(w.pool_stats(p)["live_elems"] == 0 for p in (0, 1, 2))
    assert w.get_block(17, 33, 60) == 1


def test_empty_chunk_then_edits():
    w = World()
    w.insert_dense((0, 0, 0), np.zeros(H.VOL, np.uint16))
    assert w.set_block(1, 2, 3, 8) and w.get_block(1, 2, 3) == 8
    assert w.chunk_info((0, 0, 0)).solid_count == 1
    assert w.set_block(1, 2, 3, 0) and w.get_block(1, 2, 3) == 0
    assert w.chunk_info((0, 0, 0)).root[0] == 0 and w.chunk_info((0, 0, 0)).root[1] == 0
    assert all(w.pool_stats(p)["live_elems"] == 0 for p in (0, 1, 2))


def test_edit_full_chunk_and_back():
    w = World()
    w.insert_dense((0, 0, 0), np.full(H.VOL, 1, np.uint16))
    p = (10, 20, 30)
    assert w.set_block(*p, 0) and w.get_block(*p) == 0
    assert w.get_block(11, 20, 30) == 1 and w.get_block(10, 19, 30) == 1
    assert not w.chunk_info((0, 0, 0)).root[2] & 0x80000000
    assert w.chunk_info((0, 0, 0)).solid_count == 64 ** 3 - 1
    assert w.set_block(*p, 1)
    info = w.chunk_info((0, 0, 0))
    assert info.root[2] & 0x80000000 and info.attr_uniform and info.attr_base == 1
    assert w.pool_stats(0)["live_elems"] == 0 and w.pool_stats(1)["live_elems"] == 0
    assert w.set_block(*p, 11) and w.get_block(*p) == 11 and w.get_block(10, 20, 31) == 1
    assert w.chunk_info((0, 0, 0)).has_emitter


def test_random_edits_match_dense_mirror():
    rng = np.random.default_rng(3)
    w = World()
    d = _terrain(3)
    k = (2, 0, -1)
    w.insert_dense(k, d)
    o = np.array(k) * 64
    for i in range(2000):
        x, y, z = rng.integers(0, 64, 3)
        bid = int(rng.choice([0, 0, 1, 2, 5, 7, 10, 11, 13]))
        assert w.set_block(int(o[0] + x), int(o[1] + y), int(o[2] + z), bid)
        d[H.dense_index(x, y, z)] = bid
        if i % 500 == 499:
            assert np.array_equal(w.extract_dense(k), d)
    assert np.array_equal(w.extract_dense(k), d)
    assert w.chunk_info(k).solid_count == int((d != 0).sum())
    w.remove_chunk(k)
    assert all(w.pool_stats(p)["live_elems"] == 0 for p in (0, 1, 2))


def test_mixed_full_root_has_half_dry_mask():
    # tests/full_march.rs:8-41: stone below y=32, water above -> full root, 32 dry L1 cells
    w = World()
    w.insert_dense((0, 0, 0), H.dense_from(lambda x, y, z: np.where(y < 32, 1, 13)))
    info = w.chunk_info((0, 0, 0))
    assert info.root[2] & 0x80000000 and info.has_water
    assert info.dry_mask != 2 ** 64 - 1 and info.dry_mask != 0
    assert bin(info.dry_mask).count("1") == 32


# ---- frame-level invariants on the small workload -----------------------------------------
@pytest.fixture(scope="module")
def small_world():
    return W.make_world(W.WORKLOADS[0])


def test_hiz_on_equals_hiz_off_and_no_oob(small_world):
    wl = W.WORKLOADS[0]
    cam_on = W.default_camera(wl, width=320, height=180)
    cam_off = W.default_camera(wl, flags=W.DEFAULT_FLAGS & ~A.FLAG_HIZ, width=320, height=180)
    v_on, v_off = build_view(small_world, cam_on), build_view(small_world, cam_off)
    spec, hi = W.spec_words(cam_on.flags, small_world)
    on = H.oracle_frames(small_world, v_on, spec, hi, frames=3)
    off = H.oracle_frames(small_world, v_off, spec, hi, frames=1)
    for vis, sh, hz, st in on:
        assert np.array_equal(vis, off[0][0]) and np.array_equal(sh, off[0][1])
        assert st["oob"] == 0
    assert on[1][3]["pairs_deferred"] > 0
    assert (on[0][0] != 0).mean() > 0.2 and (on[0][1] != 0).sum() > 0


def test_vis_independent_of_thread_decomposition(small_world):
    wl = W.WORKLOADS[0]
    cam = W.default_camera(wl, width=200, height=120)
    v = build_view(small_world, cam)
    spec, hi = W.spec_words(cam.flags, small_world)
    sc = O.Scene(small_world.pool(0), small_world.pool(1), small_world.pool(2), v.chunks, v.grid)
    hz = np.zeros(v.tiles[0] * v.tiles[1], np.uint32)
    a = O.render(sc, v.frame_bytes, O.spec_from(spec, hi), hz, threads=1)
    b = O.render(sc, v.frame_bytes, O.spec_from(spec, hi), hz, threads=7)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
