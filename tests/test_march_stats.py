"""The iteration heatmap and its summary (Renderer::march_stats, src/render/mod.rs:1977-2045).

The summary is checked on counters whose statistics are worked out by hand; the oracle's
per-pixel counter (march.wgsl:51-53: every marched pair adds its march_chunk iterations)
is pinned by worlds whose iteration counts are fixed by the march rules: a ray entering
a full chunk under FULL_MARCH finds its first cell solid on iteration 1, the full-root
shortcut returns before the loop (0 iterations), and every selected pair counts.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.world import World, build_view


def test_stats_of_a_ramp():
    m = A.march_stats_of(np.arange(100, dtype=np.uint32), chunks=7)
    # total 0+1+...+99; 99 pixels touched; sorted descending hot[k] = 99 - k
    assert (m["pixels"], m["iters"], m["touched"], m["capped"], m["worst"]) == (100, 4950, 99, 0, 99)
    # pct(p) = hot[((100 - p) / 100 * len) as usize % len]
    assert (m["p50"], m["p90"], m["p99"]) == (49, 89, 98)
    # top 1 % = hot[0..1] = 99; top 10 % = 99 + ... + 90 = 945
    assert m["top_1"] == pytest.approx(100.0 * 99 / 4950, rel=1e-12)
    assert m["top_10"] == pytest.approx(100.0 * 945 / 4950, rel=1e-12)
    assert m["chunks"] == 7


def test_stats_capped_and_small_counts():
    m = A.march_stats_of(np.array([600, 513, 512, 0], np.uint32))
    assert m["capped"] == 2  # strictly above MAX_ITERS = 512
    assert m["touched"] == 3 and m["worst"] == 600 and m["iters"] == 1625
    assert (m["p50"], m["p90"], m["p99"]) == (512, 600, 600)  # indices 2, 0, 0
    assert m["top_1"] == 0.0 and m["top_10"] == 0.0  # 4 // 100 = 4 // 10 = 0 pixels
    e = A.march_stats_of(np.zeros(0, np.uint32))
    assert e["pixels"] == 0 and e["p50"] == 0 and e["top_10"] == 0.0


def _full_chunks(keys):
    w = World()
    stone = np.ones(64 ** 3, np.uint16)
    for k in keys:
        w.insert_dense(k, stone)
    return w


def _heat_frame(world, cam, full_march: bool, hiz=None):
    v = build_view(world, cam)
    spec = O.spec_from(0, A.FLAG_HI_FULL_MARCH if full_march else 0)
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
    hz = np.zeros(v.tiles[0] * v.tiles[1], np.uint32) if hiz is None else hiz
    vis, _, hz, heat, st = O.render_heat(sc, v.frame_bytes, spec, hz, shadow=False)
    return v, vis, hz, heat, st


@pytest.mark.parametrize("full_march", [True, False])
def test_heat_of_one_full_chunk(full_march):
    world = _full_chunks([(0, 0, 0)])
    cam = H.camera((32.5, 32.5, -40.0), 90.0, 0.0, 48, 40, A.FLAG_HEATMAP)
    v, vis, _, heat, st = _heat_frame(world, cam, full_march)
    hit = vis != 0
    assert 0 < hit.sum() < vis.size
    # every pixel whose ray enters the box hits its first cell: 1 iteration, or the
    # shortcut's 0; rays that miss the box leave the march before the loop
    assert np.all(heat[hit] == (1 if full_march else 0))
    assert np.all(heat[~hit] == 0)
    assert int(heat.sum()) == st["iters"]


def test_heat_sums_over_selected_pairs_and_follows_hiz():
    # Two full chunks one behind the other: on a first frame (Hi-Z 0) both are selected
    # for every tile that sees them and each counts 1 iteration; with the Hi-Z words of
    # that frame the chunk behind is deferred and not recovered (its box starts beyond
    # the nearer hits), so only the front chunk counts.
    world = _full_chunks([(0, 0, 0), (0, 0, 1)])
    w, h = 32, 24
    cam = H.camera((32.5, 32.5, -40.0), 90.0, 0.0, w, h, A.FLAG_HEATMAP | A.FLAG_HIZ)
    v, vis, hz, heat, _ = _heat_frame(world, cam, True)
    hit = vis != 0
    assert 0.5 < hit.mean() < 1.0
    assert np.all(heat[~hit] == 0)
    # Independent float64 ray/box test: the rear box [0,64]x[0,64]x[64,128] is entered
    # exactly by the rays that count a second iteration (pixels within 1e-3 of its edge
    # are left out).
    f = v.frame
    fwd, right, up = (np.array(x, np.float64) for x in (f.cam_fwd, f.cam_right, f.cam_up))
    ro = np.array(f.cam_pos, np.float64)
    px, py = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
    nx = px / w * 2.0 - 1.0
    ny = 1.0 - py / h * 2.0
    d = fwd + right * (nx * f.tan_half_fov * f.aspect)[..., None] + up * (ny * f.tan_half_fov)[..., None]
    t = (64.0 - ro[2]) / d[..., 2]
    xs, ys = ro[0] + d[..., 0] * t, ro[1] + d[..., 1] * t
    margin = np.minimum(np.minimum(xs, 64.0 - xs), np.minimum(ys, 64.0 - ys)).reshape(-1)
    enters = margin > 1e-3
    misses = margin < -1e-3
    assert enters.sum() > 10 and misses.sum() > 10
    assert np.all(heat[enters] == 2) and np.all(heat[misses & hit] == 1)
    _, vis2, _, heat2, st2 = _heat_frame(world, cam, True, hiz=hz)
    assert np.array_equal(vis2, vis)
    assert np.all(heat2[hit] == 1) and st2["pairs_deferred"] > 0
