"""GPU parity on the 2048x2048x512 world (caves, sea level 96, 1 % glass): the incoherent
batch and the request server at their benchmark shapes, against the oracle.

- Incoherent batch: the 32 seeded 1280x720 views, secondary bounce on, waves on (every
  other view with reflect slope 0.35). Every view's keys,
  R16F shadow words, Hi-Z words and secondary leg records (with their sun tests) equal
  the oracle's over two frames (Hi-Z carried), the select / deferred / recover / shadow
  counts and the leg counts equal the oracle's totals, and three launch forms give the
  same bits for all 32 views.
- Server: camera requests at the four batch resolutions on the same world (chunk lists
  built on the device), each the same bits alone and in a mix of 40 others on three
  lanes and equal to the oracle; then a batch of 1000 random voxel set/clear ops is
  applied under the barrier, and requests after it equal the oracle on the edited world.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.renderer import Batch, Renderer, Server
from voxelcraft_menu_amd.world import ViewSnapshot, build_view

pytestmark = pytest.mark.gpu


def _digest(*arrays):
    h = hashlib.blake2b(digest_size=16)
    for a in arrays:
        h.update(np.ascontiguousarray(a).view(np.uint8).tobytes())
    return h.hexdigest()


def test_incoherent_batch_matches_oracle_on_the_2048_world():
    wl = W.WORKLOADS[3]
    world = W.make_world(wl)
    cams = W.random_cameras(wl, wl.views, seed=3, resolutions=[(wl.width, wl.height)])
    for i, c in enumerate(cams):  # waves on: every other view with the wave normal tilting the reflect leg
        c.anim_time, c.waves = 2.5, (0.12, 32.0, 0.32, 0.35 if i % 2 else 0.0)
    views = [build_view(world, c) for c in cams]
    specs = [W.spec_words(c.flags, world) for c in cams]
    r = Renderer(64, 64)
    r.sync_world(world)
    b = Batch(r, len(views))
    b.set_views(views, specs)
    b.set_secondary(True)
    forms = [dict(mode=A.BATCH_FLAT, graph=1, occupancy=6, variant=13),
             dict(mode=A.BATCH_PERSISTENT, graph=1, queue_chunk=2, variant=5),
             dict(mode=A.BATCH_PLAIN, graph=0)]
    digests = []
    for k, form in enumerate(forms):
        print(f"incoherent form {form}", flush=True)
        b.set_tuning(**form)
        b.reset_hiz()
        b.render()
        b.render()
        b.sync()
        dig = []
        if k == 0:
            demand, legs = b.demand(), b.secondary_count()
            tot = dict(select=0, deferred=0, recover=0, shadow_rays=0, legs=0, sun=0)
        for i, v in enumerate(views):
            gv, gs, gh, gsec = b.read_vis(i), b.read_shadow(i), b.read_hiz(i), b.read_secondary(i)
            dig.append(_digest(gv, gs, gh, gsec))
            if k:
                continue
            s, hi = specs[i]
            ref = H.oracle_frames(world, v, s, hi, frames=2)
            vis, sh, hz, st = ref[1]
            tag = f"view {i} pos={cams[i].pos}"
            assert np.array_equal(gv, vis), f"{tag}: {(gv != vis).sum()} keys differ"
            assert np.array_equal(gs, sh), f"{tag}: {(gs != sh).sum()} shadow words differ"
            assert np.array_equal(gh, hz), f"{tag}: {(gh != hz).sum()} Hi-Z words differ"
            sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
            sec, sst = O.secondary(sc, v.frame_bytes, O.spec_from(s, hi), vis)
            diff = (gsec.view(np.uint32) != sec.view(np.uint32)).any(axis=-1)
            assert not diff.any(), f"{tag}: {diff.sum()} leg records differ"
            tot["select"] += st["pairs_select"]
            tot["deferred"] += st["pairs_deferred"]
            tot["recover"] += st["pairs_recover"]
            tot["shadow_rays"] += st["rays_shadow"]
            tot["legs"] += sst["rays_secondary"]
            tot["sun"] += sst["rays_secondary_sun"]
        if k == 0:
            assert (demand["select"], demand["deferred"], demand["recover"], demand["shadow_rays"]) == (
                tot["select"], tot["deferred"], tot["recover"], tot["shadow_rays"])
            assert legs["water_legs"] + legs["glass_legs"] == tot["legs"] and legs["sun_rays"] == tot["sun"]
            assert legs["glass_legs"] > 1000 and legs["water_legs"] > 100000 and not legs["overflow"]
        digests.append(dig)
    for k in range(1, len(forms)):
        bad = [i for i in range(len(views)) if digests[k][i] != digests[0][i]]
        assert not bad, f"form {forms[k]} differs from {forms[0]} on views {bad}"
    b.close()
    r.close()


def _drain(srv, ids, limit_polls=4000):
    done = {}
    polls = 0
    while len(done) < len(ids):
        for x in srv.poll(wait_ms=50):
            done[x["id"]] = x
        polls += 1
        assert polls < limit_polls, f"only {len(done)} of {len(ids)} requests completed"
    return done


def test_server_on_the_2048_world_with_an_edit_batch():
    from voxelcraft_menu_amd import serving as S

    wl = W.WORKLOADS[4]
    world = W.make_world(wl)
    probes = W.random_cameras(wl, 6, seed=4)          # the four batch resolutions, seeded poses
    others = W.random_cameras(wl, 40, seed=41)
    r = Renderer(64, 64)
    r.sync_world(world)
    srv = Server(r, lanes=3, max_views=16, hbm_budget=2 << 30, mode=A.BATCH_FLAT, occupancy=6, variant=5)
    snap = ViewSnapshot(world)
    srv.set_chunk_table(snap.chunk_table())

    def size(c):
        return c.width, c.height

    def read(rid, c):
        return srv.read_vis(rid, *size(c)), srv.read_shadow(rid, *size(c))

    def check_oracle(c, got, tag):
        v = build_view(world, c)
        s, hi = W.spec_words(c.flags, world)
        ref = H.oracle_frames(world, v, s, hi, frames=1)[0]
        assert np.array_equal(got[0], ref[0]), f"{tag}: {(got[0] != ref[0]).sum()} keys differ from the oracle"
        assert np.array_equal(got[1], ref[1]), f"{tag}: {(got[1] != ref[1]).sum()} shadow words differ"

    alone = []
    for i, c in enumerate(probes):
        print(f"alone {i} {size(c)}", flush=True)
        srv.submit_camera(i, c, *W.spec_words(c.flags, world))
        assert _drain(srv, [i])[i]["status"] == 0
        alone.append(read(i, c))
        srv.release(i)
        check_oracle(c, alone[-1], f"probe {i} alone")
    ids = []
    for k, c in enumerate(others):
        srv.submit_camera(100 + k, c, *W.spec_words(c.flags, world))
        ids.append(100 + k)
        if k % 7 == 3 and k // 7 < len(probes):
            p = k // 7
            srv.submit_camera(1000 + p, probes[p], *W.spec_words(probes[p].flags, world))
            ids.append(1000 + p)
    print("mix submitted", flush=True)
    res = _drain(srv, ids)
    assert all(x["status"] == 0 for x in res.values())
    assert max(x["batch_views"] for x in res.values()) > 1, "nothing was batched"
    for p in range(len(probes)):
        got = read(1000 + p, probes[p])
        assert np.array_equal(got[0], alone[p][0]) and np.array_equal(got[1], alone[p][1]), f"probe {p} in the mix"
    for k in range(0, len(others), 8):
        check_oracle(others[k], read(100 + k, others[k]), f"request {100 + k} in the mix")
    for rid in ids:
        srv.release(rid)

    # one edit batch of 1000 ops, applied between frames under the barrier
    edits = S.random_edits(np.random.default_rng(44), wl, 1000)
    # 27 of them a stone block 12 blocks ahead of probe 0, so that the edit shows
    f0 = build_view(world, probes[0]).frame
    at = np.floor(np.array(f0.cam_pos) + 12.0 * np.array(f0.cam_fwd)).astype(np.int64)
    blk = np.stack(np.meshgrid(*[np.arange(-1, 2)] * 3, indexing="ij"), -1).reshape(-1, 3) + at
    edits["x"][:27], edits["y"][:27], edits["z"][:27], edits["id"][:27] = blk[:, 0], blk[:, 1], blk[:, 2], 1
    srv.barrier()
    assert world.apply_edits(edits) == 1000
    r.sync_world(world)
    snap.take(world)
    srv.set_chunk_table(snap.chunk_table())
    srv.resume()
    print("edits applied", flush=True)
    changed = 0
    for i, c in enumerate(probes):
        srv.submit_camera(2000 + i, c, *W.spec_words(c.flags, world))
    res = _drain(srv, [2000 + i for i in range(len(probes))])
    for i, c in enumerate(probes):
        assert res[2000 + i]["status"] == 0
        got = read(2000 + i, c)
        check_oracle(c, got, f"probe {i} after the edits")
        changed += int((got[0] != alone[i][0]).sum())
        srv.release(2000 + i)
    print(f"pixels changed by the edits: {changed}", flush=True)
    assert changed > 1000, "the stone block ahead of probe 0 does not show"
    srv.close()
    r.close()
