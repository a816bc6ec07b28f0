"""LOD planning and LOD chunks (src/lod.rs, tests/lod_stream.rs).

The planner tests restate the planner properties tests/lod_stream.rs checks with every
key solid: every point is covered exactly once (shares sum to 1), overlapping items
are only cross-fade partners, the cross-fade takes the pop out of a flight and the
no-fade mode holds a boundary by its hysteresis. The view builder over a draw list is
checked against the LOD-0 builder and the grid-fill rule; the march over coarse chunks
is pinned against the float64 brute-force DDA in coarse voxel units, and the fade dither
against its defining property (a fading pair draws every pixel exactly once).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.world import (GenParams, LodConfig, Planner, World, build_view, build_view_items,
                                       settle)

GEN = GenParams(seed=1337, sea_level=96, tree_density=0.0)
CAM_A = np.array([-450.0, 140.0, 1100.0], np.float32)
CAM_B = np.array([-450.0, 140.0, 300.0], np.float32)


def _probes():
    pts = []
    for x in range(-900, -99, 48):
        for z in range(100, 1301, 48):
            pts.append((x + 0.5, GEN.height(x, z) - 1.5, z + 0.5))
    return np.array(pts, np.float32)


PROBES = _probes()


def _bounds(render):
    size = (64 << render["lod"].astype(np.int64)).astype(np.float32)
    lo = render["pos"].astype(np.float32) * size[:, None]
    return lo, lo + size[:, None]


def _share(render):
    f = render["fade"]
    return np.where(f >= 0, f, 1.0 + f)


def _inside(render, pts):
    lo, hi = _bounds(render)
    return np.all((pts[:, None, :] >= lo[None]) & (pts[:, None, :] < hi[None]), axis=2)  # (P, n)


def _sweep(cfg, a, b, steps, status=None):
    p = Planner()
    out = []
    for i in range(steps):
        cam = a + (b - a) * np.float32(i / (steps - 1))
        r, _ = p.plan(tuple(float(v) for v in cam), cfg, status=status)
        out.append(r)
    return out


CASES = [("popping", LodConfig(cross_fade=False)), ("default", LodConfig()), ("band 1.0", LodConfig(fade_band=1.0)),
         ("band 3.0", LodConfig(fade_band=3.0)), ("factor 1.0", LodConfig(factor=1.0)),
         ("smootherstep", LodConfig(smootherstep=True))]


@pytest.mark.parametrize("name,cfg", CASES, ids=[c[0] for c in CASES])
def test_plan_covers_every_point_exactly_once(name, cfg):
    faded = 0
    for step, r in enumerate(_sweep(cfg, CAM_A, CAM_B, 120)):
        cov = _inside(r, PROBES).astype(np.float32) @ _share(r)
        bad = np.abs(cov - 1.0) >= 1e-4
        assert not bad.any(), f"{name} step {step}: {PROBES[bad][0]} covered {cov[bad][0]} times"
        faded += int((r["fade"] < 1.0).sum())
    if name in ("popping", "band 1.0"):
        assert faded == 0, f"{name} emitted a cross-fade"
    else:
        assert faded > 0, f"{name} never faded anything"


def test_overlaps_are_only_cross_fade_partners():
    for step, r in enumerate(_sweep(LodConfig(), CAM_A, CAM_B, 60)):
        assert len(r) < 4096
        lo, hi = _bounds(r)
        inter = np.all((lo[:, None] < hi[None]) & (hi[:, None] > lo[None]), axis=2)
        np.fill_diagonal(inter, False)
        for i, j in zip(*np.nonzero(np.triu(inter))):
            a, b = r[i], r[j]
            coarse, fine = (a, b) if a["lod"] > b["lod"] else (b, a)
            assert coarse["fade"] < 0 < fine["fade"] and coarse["fade"] == -fine["fade"], (step, a, b)
            pos, lod = fine["pos"].copy(), int(fine["lod"])
            while lod < coarse["lod"]:
                pos, lod = pos >> 1, lod + 1
            assert tuple(pos) == tuple(coarse["pos"]), (step, a, b)


def _mix_lod(r, pts):
    ins = _inside(r, pts).astype(np.float64)
    s = _share(r).astype(np.float64)
    w = ins @ s
    acc = ins @ (s * r["lod"])
    return np.where(w > 0, acc / np.maximum(w, 1e-30), np.nan)


def test_cross_fade_removes_the_pop():
    worst = []
    for cross_fade in (False, True):
        plans = _sweep(LodConfig(cross_fade=cross_fade), CAM_A, CAM_B, 300)
        prev = _mix_lod(plans[0], PROBES)
        w = 0.0
        for r in plans[1:]:
            now = _mix_lod(r, PROBES)
            ok = ~np.isnan(prev) & ~np.isnan(now)
            if ok.any():
                w = max(w, float(np.abs(now[ok] - prev[ok]).max()))
            prev = now
        worst.append(w)
    assert worst[0] > 0.99, f"popping plan never stepped a whole level ({worst[0]:.3f})"
    assert worst[1] < 0.34, f"dissolving plan stepped {worst[1]:.3f} of a level in one frame"


def _drawn_lod(r, pts):
    ins = _inside(r, pts)
    s = np.where(ins, _share(r)[None, :], -1.0)
    j = np.argmax(s, axis=1)
    return np.where(ins.any(axis=1), r["lod"][j].astype(np.int64), -1)


def test_no_fade_hysteresis_damps_boundary_flutter():
    """A camera hovering on a split boundary: the no-fade plan keeps a split key until
    the outer radius, so the drawn LOD does not flip back and forth."""
    base = np.array([-450.0, 140.0, 556.0], np.float32)
    flips = {}
    for cfg_name, cfg in (("hysteresis", LodConfig(cross_fade=False)),
                          ("no band", LodConfig(cross_fade=False, fade_band=1.0))):
        p = Planner()
        prev = None
        n = 0
        for f in range(300):
            cam = base + np.array([0.0, 0.0, np.sin(f * 0.12) * 24.0], np.float32)
            r, _ = p.plan(tuple(float(v) for v in cam), cfg)
            now = _drawn_lod(r, PROBES)
            if prev is not None:
                n += int((now != prev).sum())
            prev = now
        flips[cfg_name] = n
    assert flips["hysteresis"] < 20, flips
    assert flips["no band"] > flips["hysteresis"], flips  # the band is what holds it


def test_plan_falls_back_to_resident_coarse_chunks():
    cfg = LodConfig()
    cam = (-450.0, 140.0, 700.0)
    r_full, d_full = Planner().plan(cam, cfg)
    # No LOD-0 chunk resident: their parents draw instead, still once everywhere.
    r, d = Planner().plan(cam, cfg, status=lambda pos, lod: A.VXW_ABSENT if lod == 0 else A.VXW_SOLID)
    assert (r["lod"] > 0).all() and (r_full["lod"] == 0).any()
    cov = _inside(r, PROBES).astype(np.float32) @ _share(r)
    assert np.all(np.abs(cov - 1.0) < 1e-4)
    assert (d["lod"] == 0).any()  # the plan still asks for them
    # Nothing resident: nothing drawn, everything wanted.
    r0, d0 = Planner().plan(cam, cfg, status=lambda pos, lod: A.VXW_ABSENT)
    assert r0.size == 0 and d0.size == d_full.size


def test_coarse_stand_in_is_the_lod0_column_at_each_span_top():
    gen = GenParams(seed=7, sea_level=128, tree_density=0.0, glass_fraction=0.05, leaf_extra=0.05)
    for lod, pos in ((1, (2, 1, -1)), (2, (0, 0, 1)), (3, (-1, 0, 0))):
        s = 1 << lod
        coarse = gen.dense(pos, lod).reshape(64, 64, 64)  # [y, z, x]
        origin = [p * (64 << lod) for p in pos]
        fine_cache = {}

        def fine_block(wx, wy, wz):
            key = (wx >> 6, wy >> 6, wz >> 6)
            if key not in fine_cache:
                fine_cache[key] = gen.dense(key).reshape(64, 64, 64)
            return int(fine_cache[key][wy & 63, wz & 63, wx & 63])

        rng = np.random.default_rng(lod)
        for x, z in rng.integers(0, 64, (24, 2)):
            wx, wz = origin[0] + int(x) * s, origin[2] + int(z) * s
            ceiling = max(gen.height(wx, wz), gen.sea_level)
            for y in range(64):
                wy = origin[1] + y * s
                want = fine_block(wx, min(wy + s - 1, ceiling - 1), wz) if wy < ceiling else 0
                assert int(coarse[y, z, x]) == want, (lod, pos, x, y, z)


def _settled(cam, cfg, gen=GEN):
    w = World()
    r = settle(w, gen, Planner(), cam, cfg)
    return w, r


def test_item_view_of_lod0_chunks_is_the_lod0_view():
    w, _ = _settled((300.5, 150.0, 260.5), LodConfig(max_lod=1, view_distance=400.0))
    cam = H.camera((300.5, 150.0, 260.5), 30.0, -10.0, 160, 96, 0, sea=96.0)
    keys = w.keys()
    items = np.zeros(len(keys), A.RENDER_ITEM_DTYPE)
    items["pos"] = keys
    items["fade"] = 1.0
    a, b = build_view(w, cam), build_view_items(w, cam, items)
    assert a.n_total == b.n_total and a.n_visible == b.n_visible and a.camera_chunk == b.camera_chunk
    assert np.array_equal(a.chunks, b.chunks) and np.array_equal(a.grid, b.grid)
    assert bytes(a.frame) == bytes(b.frame)


def test_item_view_fields_and_grid_fill_order():
    cfg = LodConfig(max_lod=2, view_distance=700.0)
    pos = (300.5, 150.0, 260.5)
    w, r = _settled(pos, cfg)
    assert (r["fade"] < 1).any() and len(np.unique(r["lod"])) == 3
    cam = H.camera(pos, 30.0, -10.0, 160, 96, 0, sea=96.0)
    v = build_view_items(w, cam, r, cfg)
    ch = np.frombuffer(v.chunks.tobytes(), dtype=np.uint8).reshape(-1, 96)[:v.n_total]
    recs = [A.Chunk.from_buffer_copy(c.tobytes()) for c in ch]
    by_key = {(tuple(int(x) for x in it["pos"]), int(it["lod"])): float(it["fade"]) for it in r}
    gmin = v.frame.grid_min
    for rec in recs:
        lod = rec.lod
        size = 64 << lod
        pos_k = tuple(int(o) // size for o in rec.origin)
        assert rec.voxel_size == float(1 << lod)
        assert np.float32(by_key[(pos_k, lod)]) == np.float32(rec.fade)
    # The grid holds, per cell, the item with the largest share, then the finest LOD.
    share = lambda f: f if f >= 0 else 1.0 + f  # noqa: E731
    best = {}
    for i, rec in enumerate(recs):
        span = 1 << rec.lod
        base = [int(rec.origin[k]) // 64 - gmin[k] for k in range(3)]
        for z in range(max(base[2], 0), min(base[2] + span, 32)):
            for y in range(max(base[1], 0), min(base[1] + span, 8)):
                for x in range(max(base[0], 0), min(base[0] + span, 32)):
                    cell = x + y * 32 + z * 256
                    cand = (np.float32(share(rec.fade)), -rec.lod, i)
                    if cell not in best or cand >= best[cell]:
                        best[cell] = cand
    want = np.full(A.VX_GRID_LEN, A.VX_NO_CHUNK, np.uint32)
    for cell, (_, _, i) in best.items():
        want[cell] = i
    assert np.array_equal(v.grid, want)


# ---- the march over coarse chunks, pinned ------------------------------------------------
@pytest.mark.parametrize("lod,density", [(1, 0.01), (2, 0.05), (3, 0.02)])
def test_coarse_march_first_hit_matches_bruteforce_dda(lod, density):
    rng = np.random.default_rng(40 + lod)
    solid = rng.random((64, 64, 64)) < density
    dense = np.where(solid, 1, 0).astype(np.uint16).reshape(-1)
    w = World()
    key = (-1, 0, 2)
    w.insert_dense(key, dense, lod)
    vs = float(1 << lod)
    origin = np.array([k * (64 << lod) for k in key], np.float32)
    items = np.zeros(1, A.RENDER_ITEM_DTYPE)
    items["pos"], items["lod"], items["fade"] = key, lod, 1.0
    cam = H.camera(tuple(origin + 10.0), 0.0, 0.0, 8, 8, 0)
    v = build_view_items(w, cam, items)
    assert v.n_total == 1  # the camera sits inside the chunk
    ch = v.chunks[:1]
    rec = A.Chunk.from_buffer_copy(ch.tobytes())
    assert rec.voxel_size == vs and tuple(rec.origin) == tuple(origin)
    sc = H.scene_for(w, ch)
    fr = H.frame_for()
    sp = O.spec_from(A.FLAG_LEAF_CUTOUT | A.FLAG_GLASS, A.FLAG_HI_FULL_MARCH)
    hits = 0
    for _ in range(300):
        ro_local = rng.uniform(-20.0, 84.0, 3) if rng.random() < 0.5 else rng.uniform(0.0, 64.0, 3)
        ro = (origin + ro_local * vs).astype(np.float32)
        d = rng.normal(size=3)
        rd = (d / np.linalg.norm(d)).astype(np.float32)
        rd = np.where(np.abs(rd) < 1e-6, np.float32(1e-6), rd).astype(np.float32)
        h = H.march_chunk(sc, fr, sp, 0, ro, rd)
        ref = H.brute_dda(solid, (ro.astype(np.float64) - origin) / vs, rd.astype(np.float64))
        if ref is None:
            assert not h.hit, (ro, rd)
            continue
        hits += 1
        assert h.hit and tuple(h.voxel) == ref[0], (ro, rd, tuple(h.voxel), ref)
        if ref[1] is not None:
            assert h.normal == ref[1]
            assert abs(h.t - ref[2] * vs) <= 1e-4 * max(1.0, ref[2] * vs)
    assert hits > 20


def _twin_view(fade, frame_index=0, size=(96, 64)):
    """Two copies of one chunk at the same place, one with fade f and one with -f (the
    cross-fade pair's dither), and the same chunk alone."""
    dense = H.dense_from(lambda x, y, z: np.where((y < 20 + (x + z) % 7) | ((x - 32) ** 2 + (z - 30) ** 2 < 40),
                                                  1, 0))
    w = World()
    w.insert_dense((0, 0, 0), dense)
    w.insert_dense((3, 0, 0), dense)
    cam = H.camera((32.5, 50.0, -30.5), 90.0, -25.0, size[0], size[1], 0)
    cam.frame_index = frame_index
    items = np.zeros(2, A.RENDER_ITEM_DTYPE)
    items["pos"][0], items["pos"][1] = (0, 0, 0), (3, 0, 0)
    items["fade"] = 1.0
    v = build_view_items(w, cam, items)
    assert v.n_total == 2 and v.n_visible >= 1
    recs = [A.Chunk.from_buffer_copy(v.chunks[i].tobytes()) for i in range(2)]
    i0 = [i for i, r in enumerate(recs) if r.origin[0] == 0.0][0]
    alone = build_view_items(w, cam, items[:1])
    r1 = recs[1 - i0]
    r1.origin[0] = 0.0
    for k in range(3):
        r1.aabb_min[k], r1.aabb_max[k] = recs[i0].aabb_min[k], recs[i0].aabb_max[k]
    recs[i0].fade, r1.fade = fade, -fade
    chunks = np.frombuffer(bytes(recs[0]) + bytes(recs[1]), np.uint8).reshape(2, 96).copy()
    v.chunks = chunks
    v.frame.chunk_count = 2
    return w, v, alone, i0


@pytest.mark.parametrize("fade,frame_index", [(0.25, 0), (0.5, 3), (0.8, 17)])
def test_fade_dither_draws_each_pixel_exactly_once(fade, frame_index):
    w, v, alone, i0 = _twin_view(fade, frame_index)
    vis, _, _, _ = H.oracle_frames(w, v, 0, A.FLAG_HI_FULL_MARCH)[0]
    ref, _, _, _ = H.oracle_frames(w, alone, 0, A.FLAG_HI_FULL_MARCH)[0]
    hit = ref != 0
    assert hit.sum() > 1000
    ci_mask = np.uint64(0x1FFF << 24)
    # the same surface everywhere, drawn by one of the pair
    assert np.array_equal(vis & ~ci_mask, ref & ~ci_mask)
    ci = ((vis >> np.uint64(24)) & np.uint64(0x1FFF)).astype(np.int64)
    drawn_by_faded = (ci[hit] == i0).mean()
    assert abs(drawn_by_faded - fade) < 0.06, drawn_by_faded


@pytest.mark.parametrize("smootherstep", [False, True])
def test_fading_pairs_carry_the_schedule_of_the_parent_distance(smootherstep):
    """fine_share (src/lod.rs:158-180): a parent between size * factor and that times
    fade_band of the camera fades by the smoothstep (or smootherstep) of where in the
    band its centre lies; its children carry the complement."""
    cfg = LodConfig(smootherstep=smootherstep)
    cam = np.array([-450.0, 140.0, 700.0])
    r, _ = Planner().plan(tuple(cam), cfg)
    parents = r[r["fade"] < 0]
    assert parents.size > 10
    for it in parents:
        size = 64 << int(it["lod"])
        centre = it["pos"].astype(np.float64) * size + size / 2
        d = np.linalg.norm(centre - cam)
        inner = size * cfg.factor
        outer = inner * cfg.fade_band
        assert inner < d < outer
        x = (outer - d) / (outer - inner)
        want = x ** 3 * (x * (6 * x - 15) + 10) if smootherstep else x * x * (3 - 2 * x)
        assert abs(-float(it["fade"]) - want) < 1e-5, (it, want)
