"""GPU parity over LOD worlds: a world streamed by the LOD planner (chunks of LOD 0-2,
cross-fading pairs drawn with the per-pixel dither) rendered by the single-view paths,
by view batches (with and without the secondary bounce) and by the request server,
against the oracle, bit for bit.

The views hit coarse chunks (voxel size 2 and 4) and fading chunks of both signs, so
the march's coarse body, the dither and the grid's (share, -lod) fill all reach the
compared keys and shadow words.
"""
import numpy as np
import pytest

from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.renderer import Batch, Renderer
from voxelcraft_menu_amd.world import GenParams, LodConfig, Planner, World, build_view_items, settle

pytestmark = pytest.mark.gpu

CFG = LodConfig(max_lod=2, view_distance=800.0)
GEN = GenParams(seed=7, sea_level=128, tree_density=0.004, glass_fraction=0.02, leaf_extra=0.02)
EYE = (300.5, 170.0, 260.5)


@pytest.fixture(scope="module")
def lod_views():
    w = World()
    render = settle(w, GEN, Planner(), EYE, CFG)
    flags = W.DEFAULT_FLAGS
    poses = [(40.0, -14.0, 320, 180, flags, 0), (200.0, -30.0, 200, 120, flags, 5),
             (110.0, 5.0, 96, 64, flags & ~A.FLAG_HIZ, 63), (300.0, -45.0, 640, 360, flags, 12)]
    cams = []
    for yaw, pitch, wd, ht, fl, fi in poses:
        c = H.camera(EYE, yaw, pitch, wd, ht, fl)
        c.frame_index = fi
        cams.append(c)
    views = [build_view_items(w, c, render, CFG) for c in cams]
    specs = [W.spec_words(c.flags, w) for c in cams]
    refs = [H.oracle_frames(w, v, s, hi, frames=2) for v, (s, hi) in zip(views, specs)]
    # the fixture must reach what it is for: coarse and fading chunks in the hit keys
    lods, fading = set(), 0
    for v, ref in zip(views, refs):
        recs = [A.Chunk.from_buffer_copy(v.chunks[i].tobytes()) for i in range(v.n_total)]
        vis = ref[0][0]
        ci = ((vis[vis != 0] >> np.uint64(24)) & np.uint64(0x1FFF)).astype(np.int64)
        lods |= {recs[i].lod for i in np.unique(ci)}
        fading += sum(1 for i in ci if recs[i].fade < 1.0)
    assert lods == {0, 1, 2} and fading > 1000, (lods, fading)
    return w, views, specs, refs


@pytest.mark.parametrize("path", [A.PATH_TILES, A.PATH_PAIRS])
def test_lod_views_match_oracle(lod_views, path):
    w, views, specs, refs = lod_views
    r = Renderer(int(views[0].frame.res[0]), int(views[0].frame.res[1]), path=path)
    r.sync_world(w)
    for i, (v, (spec, hi)) in enumerate(zip(views, specs)):
        size = (int(v.frame.res[0]), int(v.frame.res[1]))
        if size != r.size:
            r.resize(*size)
        r.set_view(v)
        for f in range(2):
            tag = f"lod view{i} frame{f} path{path} res={size}"
            print(tag, flush=True)
            r.render(v.frame, spec, hi)
            r.sync()
            vis, sh, hz, st = refs[i][f]
            gv, gs, gh = r.read_vis(), r.read_shadow(), r.read_hiz()
            assert np.array_equal(gv, vis), f"{tag}: {(gv != vis).sum()} vis words differ"
            assert np.array_equal(gs, sh), f"{tag}: {(gs != sh).sum()} shadow words differ"
            assert np.array_equal(gh, hz), f"{tag}: {(gh != hz).sum()} hiz words differ"
            d = r.pair_demand()
            assert (d["select"], d["deferred"], d["recover"]) == (
                st["pairs_select"], st["pairs_deferred"], st["pairs_recover"]), tag
            assert d["shadow_rays"] == st["rays_shadow"], tag
    r.close()


@pytest.mark.parametrize("form", [dict(mode=A.BATCH_FLAT, graph=1, occupancy=6, variant=5),
                                  dict(mode=A.BATCH_PERSISTENT, graph=0, queue_chunk=2),
                                  dict(mode=A.BATCH_PLAIN, graph=0)],
                         ids=["flat65", "persistent", "plain"])
def test_lod_batch_matches_oracle(lod_views, form):
    w, views, specs, refs = lod_views
    r = Renderer(int(views[0].frame.res[0]), int(views[0].frame.res[1]))
    r.sync_world(w)
    b = Batch(r, 8)
    b.set_tuning(**form)
    b.set_views(views, specs)
    for f in range(2):
        print(f"lod batch {form} frame {f}", flush=True)
        b.render()
        b.sync()
        for i in range(len(views)):
            vis, sh, hz, _ = refs[i][f]
            tag = f"lod batch view{i} frame{f}"
            gv, gs, gh = b.read_vis(i), b.read_shadow(i), b.read_hiz(i)
            assert np.array_equal(gv, vis), f"{tag}: {(gv != vis).sum()} vis words differ"
            assert np.array_equal(gs, sh), f"{tag}: {(gs != sh).sum()} shadow words differ"
            assert np.array_equal(gh, hz), f"{tag}: {(gh != hz).sum()} hiz words differ"
    b.close()
    r.close()


@pytest.mark.parametrize("form", [dict(mode=A.BATCH_FLAT, graph=1, occupancy=6, variant=5),
                                  dict(mode=A.BATCH_PERSISTENT, graph=0)], ids=["flat65", "persistent"])
def test_lod_secondary_legs_match_oracle(form):
    """The secondary bounce over the LOD world: legs leave from hits on chunks of every
    LOD and their grid walk meets coarse and fading chunks; every leg record and sun
    test equals the oracle's."""
    from oracle import oracle as O

    w = World()
    render = settle(w, GEN, Planner(), EYE, CFG)
    cams = [H.camera(EYE, yaw, pitch, 200, 120, W.DEFAULT_FLAGS) for yaw, pitch in ((200.0, -30.0), (300.0, -45.0),
                                                                                       (0.0, -35.0))]
    for c in cams:
        c.anim_time, c.waves = 1.2, (0.12, 32.0, 0.32, 0.35)
    views = [build_view_items(w, c, render, CFG) for c in cams]
    specs = [W.spec_words(c.flags, w, W.DEFAULT_SPEC_HI) for c in cams]
    r = Renderer(64, 64)
    r.sync_world(w)
    b = Batch(r, 4)
    b.set_tuning(**form)
    b.set_views(views, specs)
    b.set_secondary(True)
    b.render()
    b.sync()
    total = sun = 0
    for i, (v, (s, hi)) in enumerate(zip(views, specs)):
        print(f"lod secondary view {i}", flush=True)
        sc = O.Scene(w.pool(0), w.pool(1), w.pool(2), v.chunks, v.grid)
        sp = O.spec_from(s, hi)
        ref_vis = O.render(sc, v.frame_bytes, sp, np.zeros(v.tiles[0] * v.tiles[1], np.uint32), shadow=False)[0]
        assert np.array_equal(b.read_vis(i), ref_vis), f"view {i} vis"
        legs, st = O.secondary(sc, v.frame_bytes, sp, ref_vis)
        diff = b.read_secondary(i).view(np.uint32) != legs.view(np.uint32)
        assert not diff.any(), f"view {i}: {diff.any(axis=-1).sum()} leg records differ"
        total += st["rays_secondary"]
        sun += st["rays_secondary_sun"]
    cnt = b.secondary_count()
    assert cnt["water_legs"] + cnt["glass_legs"] == total and cnt["sun_rays"] == sun and not cnt["overflow"]
    assert total > 1000 and sun > 500
    b.close()
    r.close()


def test_lod_views_served_match_oracle(lod_views):
    """The request server renders LOD views as the oracle's first frame of each view."""
    from voxelcraft_menu_amd.renderer import Server

    w, views, specs, refs = lod_views
    r = Renderer(64, 64)
    r.sync_world(w)
    srv = Server(r, lanes=2, max_views=4, hbm_budget=256 << 20, mode=A.BATCH_FLAT)
    for i, (v, (s, hi)) in enumerate(zip(views, specs)):
        srv.submit(i, v, s, hi)
    done = {}
    polls = 0
    while len(done) < len(views):
        for res in srv.poll(wait_ms=50):
            done[res["id"]] = res
        polls += 1
        assert polls < 2000, f"only {len(done)} of {len(views)} requests completed"
    for i, v in enumerate(views):
        assert done[i]["status"] == 0
        wd, ht = int(v.frame.res[0]), int(v.frame.res[1])
        vis, sh, _, _ = refs[i][0]  # first frame: Hi-Z words start at zero
        assert np.array_equal(srv.read_vis(i, wd, ht), vis), f"served view {i} vis"
        assert np.array_equal(srv.read_shadow(i, wd, ht), sh), f"served view {i} shadow"
        srv.release(i)
    srv.close()
    r.close()
