"""GPU parity of view batches (vx_batch_*): every launch form against the oracle and
against one-view-at-a-time rendering, bit for bit.

A batch view's keys, shadow words and Hi-Z words must equal what the oracle gives for
that view alone with the same Hi-Z input; the forms (PLAIN per-view dispatches, FLAT,
PERSISTENT queue, each with and without hipGraph replay) must agree with each other.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.renderer import Batch, Renderer
from voxelcraft_menu_amd.world import World, build_view

pytestmark = pytest.mark.gpu

FORMS = [
    dict(mode=A.BATCH_PLAIN, graph=0),
    dict(mode=A.BATCH_PLAIN, graph=1, ordered=0),
    dict(mode=A.BATCH_PLAIN, graph=0, variant=0),
    dict(mode=A.BATCH_FLAT, graph=0),
    dict(mode=A.BATCH_FLAT, graph=1, occupancy=6),
    dict(mode=A.BATCH_PERSISTENT, graph=0, queue_chunk=1),
    dict(mode=A.BATCH_PERSISTENT, graph=1, queue_chunk=3, waves_per_simd=2),
    dict(mode=A.BATCH_PERSISTENT, graph=1, variant=5),
    dict(mode=A.BATCH_PERSISTENT, graph=1, queue_chunk=4, variant=5),
    dict(mode=A.BATCH_PERSISTENT, graph=1, ordered=0, occupancy=6, queue_chunk=8),
    # screen-bin candidate lists (variant bit 8)
    dict(mode=A.BATCH_FLAT, graph=1, occupancy=6, variant=13),
    dict(mode=A.BATCH_FLAT, graph=0, variant=12),
    dict(mode=A.BATCH_PERSISTENT, graph=1, queue_chunk=2, variant=13),
]


def _form_id(f):
    return "-".join(f"{k}{v}" for k, v in f.items())


@pytest.fixture(scope="module")
def small_world():
    return W.make_world(W.WORKLOADS[0])


@pytest.fixture(scope="module")
def small_views(small_world):
    wl = W.WORKLOADS[0]
    flags = W.DEFAULT_FLAGS
    cams = W.random_cameras(wl, 5, seed=31, resolutions=[(333, 197), (64, 64), (120, 250), (40, 24)])
    h = wl.gen.height(200, 40)
    cams.append(H.camera((200.5, min(h, 127) - 0.5, 40.5), 120.0, 10.0, 96, 72, flags | A.FLAG_UNDERWATER))
    cams.append(H.camera((60.5, 120.3, 70.5), 30.0, -5.0, 72, 40, flags))
    cams.append(H.camera((128.5, 150.0, 128.5), 0.0, 60.0, 80, 48, flags & ~A.FLAG_HIZ))
    views = [build_view(small_world, c) for c in cams]
    specs = [W.spec_words(c.flags, small_world) for c in cams]
    refs = [H.oracle_frames(small_world, v, s, hi, frames=2) for v, (s, hi) in zip(views, specs)]
    return views, specs, refs


def _renderer(world, views):
    r = Renderer(int(views[0].frame.res[0]), int(views[0].frame.res[1]))
    r.sync_world(world)
    return r


@pytest.mark.parametrize("form", FORMS, ids=_form_id)
def test_batch_forms_match_oracle(small_world, small_views, form):
    views, specs, refs = small_views
    r = _renderer(small_world, views)
    b = Batch(r, 16)
    b.set_tuning(**form)
    b.set_views(views, specs)
    for f in range(2):
        print(f"form {_form_id(form)} frame {f}", flush=True)
        b.render()
        b.sync()
        tot = {"select": 0, "deferred": 0, "recover": 0, "shadow_rays": 0}
        for i in range(len(views)):
            vis, sh, hz, st = refs[i][f]
            tag = f"{_form_id(form)} view{i} frame{f} res={tuple(views[i].frame.res)}"
            gv, gs, gh = b.read_vis(i), b.read_shadow(i), b.read_hiz(i)
            assert np.array_equal(gv, vis), f"{tag}: {(gv != vis).sum()} vis words differ"
            assert np.array_equal(gs, sh), f"{tag}: {(gs != sh).sum()} shadow words differ"
            assert np.array_equal(gh, hz), f"{tag}: {(gh != hz).sum()} hiz words differ"
            tot["select"] += st["pairs_select"]
            tot["deferred"] += st["pairs_deferred"]
            tot["recover"] += st["pairs_recover"]
            tot["shadow_rays"] += st["rays_shadow"]
        d = b.demand()
        assert {k: d[k] for k in tot} == tot, (d, tot)
    b.close()
    r.close()


def test_graph_replays_stay_exact_with_eager_work_between(small_world, small_views):
    """Every replay of a captured batch equals the eager result, with single-view eager
    renders and synchronisations on the same stream between the replays."""
    views, specs, refs = small_views
    r = _renderer(small_world, views)
    b = Batch(r, 16)
    b.set_tuning(mode=A.BATCH_PERSISTENT, graph=1, queue_chunk=2)
    b.set_views(views, specs)
    v0 = views[0]
    r.resize(int(v0.frame.res[0]), int(v0.frame.res[1]))
    r.set_view(v0)
    for rep in range(5):
        b.render()
        r.render(v0.frame, *specs[0])
        b.sync()
        f = min(rep, 1)  # Hi-Z reaches its steady state after the first frame
        for i in range(len(views)):
            assert np.array_equal(b.read_vis(i), refs[i][f][0]), f"replay {rep} view {i}"
            assert np.array_equal(b.read_shadow(i), refs[i][f][1]), f"replay {rep} view {i}"
        assert np.array_equal(r.read_vis(), refs[0][0][0])
    b.close()
    r.close()


def test_new_view_set_and_pool_edits_follow_into_the_graph():
    """A captured persistent batch replays correctly after set_views with a different
    view set and after edits change the pools (the graph is re-captured when a pool
    moves or changes length)."""
    wl = W.WORKLOADS[0]
    world = W.make_world(wl)
    cams_a = W.random_cameras(wl, 4, seed=41, resolutions=[(96, 64), (160, 90)])
    cams_b = W.random_cameras(wl, 6, seed=42, resolutions=[(64, 48), (200, 120)])
    r = Renderer(96, 64)
    r.sync_world(world)
    b = Batch(r, 8)
    b.set_tuning(mode=A.BATCH_PERSISTENT, graph=1)
    rng = np.random.default_rng(3)
    for step, cams in enumerate([cams_a, cams_b, cams_a]):
        if step == 2:
            e = np.zeros(3000, A.EDIT_DTYPE)
            e["x"] = rng.integers(0, 256, 3000)
            e["y"] = rng.integers(64, 192, 3000)
            e["z"] = rng.integers(0, 256, 3000)
            e["id"] = rng.choice([0, 1, 2, 8, 10, 11], 3000)
            assert world.apply_edits(e) == 3000
            r.sync_world(world)
        views = [build_view(world, c) for c in cams]
        specs = [W.spec_words(c.flags, world) for c in cams]
        b.set_views(views, specs)
        b.render()
        b.sync()
        for i, (v, (s, hi)) in enumerate(zip(views, specs)):
            ref = H.oracle_frames(world, v, s, hi, frames=1)[0]
            assert np.array_equal(b.read_vis(i), ref[0]), f"step {step} view {i}"
            assert np.array_equal(b.read_shadow(i), ref[1]), f"step {step} view {i}"
    b.close()
    r.close()


def test_config_batch_64_views_all_forms_identical():
    """The 64-view batch on the 1024x1024x256 world: batched forms and per-view plain
    dispatches give identical bits for every view over two frames, and a sample of
    views (reduced to the two smallest sizes) matches the oracle."""
    wl = W.WORKLOADS[2]
    world = W.make_world(wl)
    cams = W.random_cameras(wl, wl.views, seed=2)
    views = [build_view(world, c) for c in cams]
    specs = [W.spec_words(c.flags, world) for c in cams]
    r = Renderer(256, 144)
    r.sync_world(world)
    b = Batch(r, 64)
    b.set_views(views, specs)
    results = {}
    for form in (dict(mode=A.BATCH_PLAIN, graph=0), dict(mode=A.BATCH_FLAT, graph=1),
                 dict(mode=A.BATCH_PERSISTENT, graph=1, queue_chunk=2),
                 dict(mode=A.BATCH_FLAT, graph=1, occupancy=6, variant=13)):
        print(f"64-view form {_form_id(form)}", flush=True)
        b.set_tuning(**form)
        b.reset_hiz()
        b.render()
        b.render()
        b.sync()
        results[_form_id(form)] = [(b.read_vis(i), b.read_shadow(i), b.read_hiz(i)) for i in range(len(views))]
    keys = list(results)
    for k in keys[1:]:
        for i in range(len(views)):
            for j, name in enumerate(("vis", "shadow", "hiz")):
                assert np.array_equal(results[k][i][j], results[keys[0]][i][j]), f"{k} view {i} {name}"
    small = [i for i, v in enumerate(views) if int(v.frame.res[0]) <= 640][:4]
    large = [next(i for i, v in enumerate(views) if int(v.frame.res[0]) == w) for w in (1280, 1920)]
    for i in small + large:
        s, hi = specs[i]
        ref = H.oracle_frames(world, views[i], s, hi, frames=2)
        assert np.array_equal(results[keys[0]][i][0], ref[1][0]), f"view {i} vs oracle"
        assert np.array_equal(results[keys[0]][i][1], ref[1][1]), f"view {i} vs oracle"
    b.close()
    r.close()


def test_screen_bins_change_no_bit():
    """Screen-bin candidate lists only drop chunks that every test of the bin's tiles
    rejects: keys, shadow words, Hi-Z words and the select / deferred / recover counts
    equal the whole-list forms, for cameras at the edges of the bin model (looking
    along an axis, straight up and down, wide and narrow fields of view, jitter, odd
    sizes with partial tiles, inside the terrain, under water), and the oracle for a
    sample of them."""
    wl = W.WORKLOADS[0]
    world = W.make_world(wl)
    flags = W.DEFAULT_FLAGS
    from voxelcraft_menu_amd.world import CameraParams
    import math
    cams = W.random_cameras(wl, 6, seed=97, resolutions=[(333, 197), (130, 250), (64, 40)])

    def cam(pos, yaw, pitch, w, h, fl=flags, **kw):
        return CameraParams(pos=pos, yaw=math.radians(yaw), pitch=math.radians(pitch), width=w, height=h,
                            flags=fl, **kw)

    cams += [
        cam((128.5, 140.0, 128.5), 0.0, 0.0, 200, 120),              # along +x, horizon row on tile rows
        cam((128.0, 140.0, 128.0), 90.0, 0.0, 128, 128),             # along +z, on chunk corners
        cam((100.5, 200.0, 90.5), 10.0, 89.9, 96, 64),               # straight up
        cam((100.5, 200.0, 90.5), 30.0, -89.9, 96, 64),              # straight down
        cam((60.5, 150.0, 200.5), 45.0, -20.0, 160, 90, fov_deg=150.0),
        cam((60.5, 150.0, 200.5), 200.0, -10.0, 160, 90, fov_deg=8.0),
        cam((30.5, 135.0, 30.5), 300.0, -15.0, 150, 101, jitter=(0.37, -0.42)),
        cam((120.5, 60.0, 130.5), 120.0, 5.0, 72, 40),               # inside the terrain
    ]
    h = wl.gen.height(200, 40)
    cams.append(H.camera((200.5, min(h, 127) - 0.5, 40.5), 120.0, 10.0, 96, 72, flags | A.FLAG_UNDERWATER))
    views = [build_view(world, c) for c in cams]
    specs = [W.spec_words(c.flags, world) for c in cams]
    r = _renderer(world, views)
    b = Batch(r, len(views))
    b.set_views(views, specs)
    out = {}
    for var in (5, 13, 4, 12):
        for mode in (A.BATCH_FLAT, A.BATCH_PERSISTENT):
            form = dict(mode=mode, graph=1, occupancy=6, variant=var)
            print(f"bins form {_form_id(form)}", flush=True)
            b.set_tuning(**form)
            b.reset_hiz()
            got = []
            for f in range(2):
                b.render()
                b.sync()
                got.append(([(b.read_vis(i), b.read_shadow(i), b.read_hiz(i)) for i in range(len(views))],
                            b.demand()))
            out[_form_id(form)] = got
    keys = list(out)
    for k in keys[1:]:
        for f in range(2):
            assert out[k][f][1] == out[keys[0]][f][1], f"{k} frame {f}: demand {out[k][f][1]}"
            for i in range(len(views)):
                for j, name in enumerate(("vis", "shadow", "hiz")):
                    assert np.array_equal(out[k][f][0][i][j], out[keys[0]][f][0][i][j]), f"{k} frame {f} view {i} {name}"
    for i in (6, 8, 10, 12):
        s, hi = specs[i]
        ref = H.oracle_frames(world, views[i], s, hi, frames=2)
        got = out["mode1-graph1-occupancy6-variant13"]
        for f in range(2):
            for j, name in enumerate(("vis", "shadow", "hiz")):
                assert np.array_equal(got[f][0][i][j], ref[f][j]), f"view {i} frame {f} {name} vs oracle"
    b.close()
    r.close()


def _sec_world():
    from tests.test_secondary_oracle import _dense
    world = World()
    world.insert_dense((0, 0, 0), _dense())
    return world


@pytest.mark.parametrize("form", [dict(mode=A.BATCH_FLAT, graph=0), dict(mode=A.BATCH_PERSISTENT, graph=1),
                                  dict(mode=A.BATCH_PLAIN, graph=0)], ids=_form_id)
def test_secondary_legs_match_oracle(form):
    """Secondary legs (water refract/reflect, glass transmit/reflect, and the water legs
    behind glass) of every pixel, with the sun test at each hit, equal the oracle's records
    bit for bit: on the constructed water/glass chunk (sea level at its water surface, both
    water sun-ray ranges, flat and waved water) and on views of the 1024x1024x256 world
    with waves on (reflect slope 0 and 0.35, two animation times)."""
    flags = A.FLAG_WATER_REFLECT | A.FLAG_WATER_REFRACT | A.FLAG_GLASS | A.FLAG_SHADOWS | A.FLAG_HIZ
    cases = []
    w1 = _sec_world()
    waved = H.camera((30.5, 47.3, 6.5), 90.0, -25.0, 96, 64, flags | 512 | 1024 | 2048 | 4096, sea=32.0)
    waved.anim_time, waved.waves = 1.7, (0.12, 32.0, 0.32, 0.35)
    cases.append((w1, [H.camera((30.5, 47.3, 6.5), 90.0, -25.0, 96, 64, flags, sea=32.0),
                       H.camera((30.5, 47.3, 6.5), 90.0, -25.0, 96, 64, flags | A.FLAG_WATER_SHADOW_CUT, sea=32.0),
                       waved],
                  A.FLAG_HI_FULL_MARCH | A.FLAG_HI_GLASS_REFLECT))
    wl = W.WORKLOADS[1]
    w2 = W.make_world(wl)
    cams = W.random_cameras(wl, 4, seed=91, resolutions=[(320, 180), (200, 120)])
    for i, c in enumerate(cams):
        c.anim_time, c.waves = 0.5 + i, (0.12, 32.0, 0.32, 0.35 if i % 2 else 0.0)
    cases.append((w2, cams, W.DEFAULT_SPEC_HI))
    for world, cams, spec_hi in cases:
        views = [build_view(world, c) for c in cams]
        specs = [W.spec_words(c.flags, world, spec_hi) for c in cams]
        r = Renderer(64, 64)
        r.sync_world(world)
        b = Batch(r, 8)
        b.set_tuning(**form)
        b.set_views(views, specs)
        b.set_secondary(True)
        refs = []
        for i, (v, (s, hi)) in enumerate(zip(views, specs)):
            sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
            sp = O.spec_from(s, hi)
            ref_vis = O.render(sc, v.frame_bytes, sp, np.zeros(v.tiles[0] * v.tiles[1], np.uint32), shadow=False)[0]
            refs.append((ref_vis,) + O.secondary(sc, v.frame_bytes, sp, ref_vis))
        for frame in range(2):  # the second frame clears the first one's records, then writes its own
            b.render()
            b.sync()
            total = sun = 0
            for i in range(len(views)):
                ref_vis, legs, st = refs[i]
                assert np.array_equal(b.read_vis(i), ref_vis)
                got = b.read_secondary(i)
                diff = got.view(np.uint32) != legs.view(np.uint32)
                assert not diff.any(), f"frame {frame} view {i}: {diff.any(axis=-1).sum()} leg records differ"
                total += st["rays_secondary"]
                sun += st["rays_secondary_sun"]
        cnt = b.secondary_count()
        assert cnt["water_legs"] + cnt["glass_legs"] == total and not cnt["overflow"]
        assert cnt["sun_rays"] == sun
        assert total > 100 and sun > 10
        b.close()
        r.close()


def test_secondary_refuses_water_look(small_world):
    """The water-look ripples are not carried into the legs: such views are refused."""
    wl = W.WORKLOADS[0]
    cam = W.random_cameras(wl, 1, seed=3, resolutions=[(64, 48)])[0]
    v = build_view(small_world, cam)
    r = Renderer(64, 48)
    r.sync_world(small_world)
    b = Batch(r, 2)
    b.set_views([v], [W.spec_words(cam.flags, small_world, W.DEFAULT_SPEC_HI | A.FLAG_HI_WATER_LOOK)])
    with pytest.raises(RuntimeError):
        b.set_secondary(True)
    b.close()
    r.close()


def test_screen_bins_follow_view_set_changes(small_world, monkeypatch):
    """The bin lists are rebuilt for each view set: a second set with the same class
    ranges replays the captured graph over rebuilt lists; a set whose lists exceed the
    size limit (lowered here through VX_BIN_LIST_MAX_BYTES) drops to whole lists and the
    graph with it; a set from cameras (lists built on the device) uses bins too. Every
    frame equals the whole-list form (variant 5) on the same set."""
    from voxelcraft_menu_amd.world import ViewSnapshot

    wl = W.WORKLOADS[0]
    sets = [W.random_cameras(wl, 6, seed=s, resolutions=[(200, 120), (96, 64)]) for s in (301, 302)]
    for c, a in zip(sets[1], sets[0]):
        c.width, c.height = a.width, a.height  # same sizes: the same FLAT class ranges
    views = [[build_view(small_world, c) for c in cams] for cams in sets]
    specs = [[W.spec_words(c.flags, small_world) for c in cams] for cams in sets]
    r = _renderer(small_world, views[0])
    ref_b = Batch(r, 8)
    ref_b.set_tuning(mode=A.BATCH_FLAT, graph=0, occupancy=6, variant=5)
    b = Batch(r, 8)
    b.set_tuning(mode=A.BATCH_FLAT, graph=1, occupancy=6, variant=13)

    def frames(batch, k):
        batch.reset_hiz()
        batch.render()
        batch.render()
        batch.sync()
        return [(batch.read_vis(i), batch.read_shadow(i), batch.read_hiz(i)) for i in range(len(views[k]))], \
            batch.demand()

    refs = []
    for k in range(2):
        ref_b.set_views(views[k], specs[k])
        refs.append(frames(ref_b, k))

    def check(k, tag):
        got = frames(b, k)
        assert got[1] == refs[k][1], f"{tag}: demand {got[1]} != {refs[k][1]}"
        for i in range(len(views[k])):
            for j, name in enumerate(("vis", "shadow", "hiz")):
                assert np.array_equal(got[0][i][j], refs[k][0][i][j]), f"{tag}: view {i} {name}"

    b.set_views(views[0], specs[0])
    check(0, "first set")
    b.set_views(views[1], specs[1])
    check(1, "second set, graph kept")
    monkeypatch.setenv("VX_BIN_LIST_MAX_BYTES", "64")
    b.set_views(views[0], specs[0])
    check(0, "lists over the limit: whole lists")
    monkeypatch.delenv("VX_BIN_LIST_MAX_BYTES")
    b.set_views(views[1], specs[1])
    check(1, "bins again")
    snap = ViewSnapshot(small_world)
    b.set_chunk_table(snap.chunk_table())
    b.set_cameras(sets[0], specs[0])
    check(0, "device-built lists with bins")
    b.close()
    ref_b.close()
    r.close()


def test_moved_cameras_carry_hiz_and_match_oracle(small_world):
    """vx_batch_move_cameras: the views' cameras turn by 0.01 rad of yaw per frame (as the
    reference's bench_frames turns its camera); each frame's keys, shadow words and Hi-Z
    words equal the oracle's for the moved camera with the previous frame's Hi-Z words,
    and a set whose sizes differ is refused."""
    import dataclasses

    from voxelcraft_menu_amd.world import ViewSnapshot

    wl = W.WORKLOADS[0]
    cams = W.random_cameras(wl, 5, seed=77, resolutions=[(200, 120), (96, 64), (64, 40)])
    specs = [W.spec_words(c.flags, small_world) for c in cams]
    r = _renderer(small_world, [build_view(small_world, cams[0])])
    b = Batch(r, 8)
    b.set_tuning(mode=A.BATCH_FLAT, graph=1, occupancy=6, variant=13)
    b.set_chunk_table(ViewSnapshot(small_world).chunk_table())
    hz = [np.zeros(((c.width + 7) // 8) * ((c.height + 7) // 8), np.uint32) for c in cams]
    recovered = 0
    for k in range(4):
        moved = [dataclasses.replace(c, yaw=c.yaw + 0.01 * k) for c in cams]
        b.set_cameras(moved, specs, moved=k > 0)
        b.render()
        b.sync()
        for i, c in enumerate(moved):
            v = build_view(small_world, c)
            s, hi = specs[i]
            vis, sh, hz[i], st = H.oracle_frames(small_world, v, s, hi, frames=1, hiz=hz[i])[0]
            recovered += st["pairs_recover"]
            tag = f"step {k} view {i}"
            assert np.array_equal(b.read_vis(i), vis), tag
            assert np.array_equal(b.read_shadow(i), sh), tag
            assert np.array_equal(b.read_hiz(i), hz[i]), tag
    print(f"recovered pairs over the turning frames: {recovered}", flush=True)
    bad = [dataclasses.replace(c, width=c.width + 8) for c in cams]
    with pytest.raises(RuntimeError):
        b.set_cameras(bad, specs, moved=True)
    b.close()
    r.close()
