"""GPU parity over LOD worlds: a world streamed by the LOD planner (chunks of LOD 0-2,
cross-fading pairs drawn with the per-pixel dither) rendered by the single-view paths
and by view batches, against the oracle, bit for bit.

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
