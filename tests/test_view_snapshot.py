"""View snapshots (vxw_snapshot_*): the chunk list of a view set up from a snapshot is
the one vxw_build_view gave when the snapshot was taken, however the world has been
edited since, and a new take follows the edits."""
import numpy as np

from voxelcraft_menu_amd import serving as S
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.world import ViewSnapshot, build_view


def _same(a, b):
    return (a.n_total == b.n_total and a.n_visible == b.n_visible and a.camera_chunk == b.camera_chunk
            and np.array_equal(a.chunks, b.chunks) and np.array_equal(a.grid, b.grid) and bytes(a.frame) == bytes(b.frame))


def test_snapshot_views_are_the_world_at_take_time():
    wl = W.WORKLOADS[0]
    world = W.make_world(wl)
    cams = W.random_cameras(wl, 5, seed=9, resolutions=[(160, 96), (64, 64)])
    snap = ViewSnapshot(world)
    assert snap.any_emitter_resident() == world.any_emitter_resident()
    before = [build_view(world, c) for c in cams]
    world.apply_edits(S.random_edits(np.random.default_rng(1), wl, 3000))
    after = [build_view(world, c) for c in cams]
    assert sum(not _same(a, b) for a, b in zip(after, before)) >= 3  # the edits reach the views
    assert all(_same(snap.build_view(c), b) for c, b in zip(cams, before))
    snap.take(world)
    assert all(_same(snap.build_view(c), a) for c, a in zip(cams, after))
    snap.close()
