"""The per-chunk block layout of the pools (vxw_flat_layout, vx_pool_set_flat).

CPU: the oracle renders the same keys, Hi-Z words and shadow words from the block layout
as from the reference pools (every lookup of common.wgsl:266-356 returns the same
block), on the 640x360 workload and on a world with water, glass, leaves and glowstone;
every block starts on a 128-byte line and the bytes of a chunk's runs are its own.
GPU: the tile march and the batched forms read the block layout and give the bits of
the reference layout."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.world import ViewSnapshot, World, build_view, flat_layout, relocate_view


def _table(world):
    snap = ViewSnapshot(world)
    t = snap.chunk_table()
    snap.close()
    return t


def _render_both(world, view, spec, hi, frames=2):
    table = _table(world)
    flat, moved = flat_layout(world.pool(0), world.pool(1), world.pool(2), table)
    fv = relocate_view(view, table, moved)
    sa = O.Scene(world.pool(0), world.pool(1), world.pool(2), view.chunks, view.grid)
    sb = O.Scene(flat, flat, flat, fv.chunks, fv.grid)
    ha = np.zeros(view.tiles[0] * view.tiles[1], np.uint32)
    hb = ha.copy()
    out = []
    for _ in range(frames):
        va, sha, ha, _ = O.render(sa, view.frame_bytes, O.spec_from(spec, hi), ha)
        vb, shb, hb, _ = O.render(sb, fv.frame_bytes, O.spec_from(spec, hi), hb)
        out.append((va, vb, sha, shb, ha.copy(), hb.copy()))
    return table, flat, moved, out


def test_flat_layout_renders_the_reference_bits_on_the_small_workload():
    wl = W.WORKLOADS[0]
    world = W.make_world(wl)
    cam = W.default_camera(wl)
    view = build_view(world, cam)
    spec, hi = W.spec_words(cam.flags, world)
    table, flat, moved, frames = _render_both(world, view, spec, hi)
    for va, vb, sha, shb, ha, hb in frames:
        assert (va != 0).sum() > 1000
        assert np.array_equal(va, vb) and np.array_equal(sha, shb) and np.array_equal(ha, hb)
    # Blocks start on 128-byte lines: run pointers in 16-byte units are multiples of 8.
    rec = moved.view(np.uint32).reshape(-1, 24)
    full = (rec[:, 14] & 0x80000000) != 0
    nonempty = (rec[:, 12] | rec[:, 13]) != 0
    ptr = rec[:, 14] & 0x7FFFFFFF
    sel = ~full & nonempty
    assert sel.any() and np.all(ptr[sel] % 8 == 0)
    assert flat.size % 32 == 0
    # Every non-uniform chunk's attribute table lies inside its own block.
    uni = (rec[:, 17] & 1) != 0
    starts = np.sort(ptr[sel] * 4)
    assert np.all(rec[~uni & sel, 16] > ptr[~uni & sel] * 4)
    assert len(np.unique(starts)) == len(starts)


def test_flat_layout_with_water_glass_leaves_and_emitters():
    wl = W.WORKLOADS[0]
    gen = dataclasses.replace(wl.gen, glass_fraction=0.05, leaf_extra=0.05, cave_fraction=0.5)
    wl = dataclasses.replace(wl, gen=gen)
    world = World()
    world.generate(gen)
    # glowstone blocks on the surface so the emitter branch of the sun ray runs
    rng = np.random.default_rng(5)
    for _ in range(200):
        x, z = int(rng.integers(0, 256)), int(rng.integers(0, 256))
        world.set_block(x, int(gen.height(x, z)), z, 11)
    cam = W.default_camera(wl)
    view = build_view(world, cam)
    spec, hi = W.spec_words(cam.flags, world)
    _, _, _, frames = _render_both(world, view, spec, hi, frames=1)
    va, vb, sha, shb, ha, hb = frames[0]
    assert (va != 0).sum() > 1000 and (sha != 0).sum() > 100
    assert np.array_equal(va, vb) and np.array_equal(sha, shb) and np.array_equal(ha, hb)


def test_flat_layout_refuses_light_tables_and_short_pools():
    wl = W.WORKLOADS[0]
    world = W.make_world(wl)
    table = _table(world).copy()
    rec = table.view(np.uint32).reshape(-1, 24)
    rec[0, 19] = 0  # light table present
    with pytest.raises(Exception):
        flat_layout(world.pool(0), world.pool(1), world.pool(2), table)
    with pytest.raises(Exception):
        flat_layout(world.pool(0)[:4], world.pool(1), world.pool(2), _table(world))


@pytest.mark.gpu
def test_gpu_flat_layout_single_view_and_batch_match_the_reference_layout():
    from voxelcraft_menu_amd import _abi as A
    from voxelcraft_menu_amd.renderer import Batch, Renderer

    wl = W.WORKLOADS[0]
    world = W.make_world(wl)
    cam = W.default_camera(wl)
    view = build_view(world, cam)
    spec, hi = W.spec_words(cam.flags, world)
    table = _table(world)
    flat, moved = flat_layout(world.pool(0), world.pool(1), world.pool(2), table)
    fv = relocate_view(view, table, moved)
    print(f"[flat layout] single view {cam.width}x{cam.height}, {flat.nbytes} block bytes", flush=True)
    r = Renderer(cam.width, cam.height, path=A.PATH_TILES, device=0)
    r.sync_world(world)
    out = {}
    for name, v, words in (("reference", view, None), ("flat", fv, flat)):
        r.set_flat_pools(words)
        r.set_view(v)
        r.write_hiz(np.zeros(view.tiles[0] * view.tiles[1], np.uint32))
        frames = []
        for _ in range(2):
            r.render(v.frame, spec, hi)
            r.sync()
            frames.append((r.read_vis(), r.read_shadow(), r.read_hiz()))
        out[name] = frames
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), view.chunks, view.grid)
    hz = np.zeros(view.tiles[0] * view.tiles[1], np.uint32)
    for k in range(2):
        vis, sh, hz, _ = O.render(sc, view.frame_bytes, O.spec_from(spec, hi), hz)
        for name in out:
            gv, gs, gh = out[name][k]
            assert np.array_equal(gv, vis) and np.array_equal(gs, sh) and np.array_equal(gh, hz), (name, k)

    # Batched forms over views of the same world, reference pools then the block layout.
    cams = W.random_cameras(wl, 6, seed=11, resolutions=[(256, 144), (320, 200), (640, 360)])
    views = [build_view(world, c) for c in cams]
    specs = [W.spec_words(c.flags, world) for c in cams]
    fviews = [relocate_view(v, table, moved) for v in views]
    b = Batch(r, len(views))
    res = {}
    for name, vs_, words in (("reference", views, None), ("flat", fviews, flat)):
        r.set_flat_pools(words)
        b.set_views(vs_, specs)
        for mode in (A.BATCH_FLAT, A.BATCH_PERSISTENT):
            b.set_tuning(mode=mode, occupancy=6, variant=13, ordered=1, queue_chunk=2, waves_per_simd=4, graph=1)
            b.reset_hiz()
            for _ in range(2):
                b.render()
            b.sync()
            res[(name, mode)] = [(b.read_vis(i), b.read_shadow(i)) for i in range(len(views))]
    b.close()
    r.set_flat_pools(None)
    r.close()
    for mode in (A.BATCH_FLAT, A.BATCH_PERSISTENT):
        for i in range(len(views)):
            va, sa = res[("reference", mode)][i]
            vb, sb = res[("flat", mode)][i]
            assert (va != 0).sum() > 0
            assert np.array_equal(va, vb) and np.array_equal(sa, sb), (mode, i)
