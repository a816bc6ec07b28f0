"""Chunk lists built on the device (vx_batch_set_chunk_table / vx_batch_set_cameras):
for every camera, the list, grid, visible count and camera chunk are the bytes the host
builder gives (build_chunk_list, src/render/mod.rs:1447-1564), on the 8192-chunk world
and after edits, and a batch set from cameras renders the same bits as one set from host
lists."""
import numpy as np
import pytest

from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd import serving as S
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.renderer import Batch, Renderer
from voxelcraft_menu_amd.world import ViewSnapshot, build_view

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big():
    wl = W.WORKLOADS[4]
    world = W.make_world(wl)
    flags = W.DEFAULT_FLAGS
    cams = W.random_cameras(wl, 20, seed=71, resolutions=[(320, 180), (200, 120), (96, 64), (64, 40)])
    h = wl.gen.height(700, 900)
    cams.append(H.camera((700.5, min(h, wl.gen.sea_level) - 2.5, 900.5), 30.0, 5.0, 96, 64,
                         flags | A.FLAG_UNDERWATER, sea=float(wl.gen.sea_level)))
    cams.append(H.camera((1000.5, 60.0, 1000.5), 120.0, -5.0, 96, 64, flags))  # inside the terrain
    cams.append(H.camera((10.5, 300.0, 10.5), 45.0, 70.0, 96, 64, flags & ~A.FLAG_HIZ))  # corner, looking up
    return wl, world, cams


def _check_lists(b, world, cams):
    for i, cam in enumerate(cams):
        v = build_view(world, cam)
        nt, nv, cc = b.view_counts(i)
        ch, grid = b.read_chunks(i)
        tag = f"view {i} pos={cam.pos}"
        assert (nt, nv, cc) == (v.n_total, v.n_visible, v.camera_chunk), tag
        assert np.array_equal(ch, v.chunks[:v.n_total]), f"{tag}: chunk records differ"
        assert np.array_equal(grid, v.grid), f"{tag}: grid differs"


def test_device_lists_are_the_host_lists(big):
    wl, world, cams = big
    snap = ViewSnapshot(world)
    table = snap.chunk_table()
    assert 1000 < len(table) <= A.VX_MAX_CHUNKS + 1
    r = Renderer(64, 64)
    r.sync_world(world)
    b = Batch(r, 32)
    b.set_chunk_table(table)
    specs = [W.spec_words(c.flags, world) for c in cams]
    print(f"{len(cams)} cameras over {len(table)} chunks", flush=True)
    b.set_cameras(cams, specs)
    _check_lists(b, world, cams)
    # after an edit batch: a new table, the same cameras
    world.apply_edits(S.random_edits(np.random.default_rng(3), wl, 4000))
    snap.take(world)
    r.sync_world(world)
    b.set_chunk_table(snap.chunk_table())
    b.set_cameras(cams, specs)
    _check_lists(b, world, cams)
    b.close()
    r.close()


def test_camera_views_render_as_host_views(big):
    wl, world, cams = big
    cams = cams[:12] + cams[-3:]
    specs = [W.spec_words(c.flags, world) for c in cams]
    snap = ViewSnapshot(world)
    r = Renderer(64, 64)
    r.sync_world(world)
    host = Batch(r, 16)
    host.set_tuning(mode=A.BATCH_FLAT, graph=1, occupancy=6, variant=5)
    host.set_views([build_view(world, c) for c in cams], specs)
    dev = Batch(r, 16)
    dev.set_tuning(mode=A.BATCH_FLAT, graph=1, occupancy=6, variant=5)
    dev.set_chunk_table(snap.chunk_table())
    dev.set_cameras(cams, specs)
    for f in range(2):
        print(f"frame {f}", flush=True)
        host.render()
        host.sync()
        dev.render()
        dev.sync()
        for i in range(len(cams)):
            tag = f"frame {f} view {i}"
            assert np.array_equal(dev.read_vis(i), host.read_vis(i)), tag
            assert np.array_equal(dev.read_shadow(i), host.read_shadow(i)), tag
            assert np.array_equal(dev.read_hiz(i), host.read_hiz(i)), tag
    dev.close()
    host.close()
    r.close()


def test_chunk_table_limits():
    r = Renderer(64, 64)
    b = Batch(r, 2)
    with pytest.raises(RuntimeError):
        b.set_chunk_table(np.zeros((A.VX_CHUNK_TABLE_MAX + 1, 96), np.uint8))
    lod1 = np.zeros((1, 96), np.uint8)
    rec = A.Chunk.from_buffer_copy(lod1.tobytes())
    rec.lod, rec.voxel_size, rec.fade = 1, 2.0, 1.0
    with pytest.raises(RuntimeError):
        b.set_chunk_table(np.frombuffer(bytes(rec), np.uint8).reshape(1, 96))
    # an empty table gives empty lists
    b.set_chunk_table(np.zeros((0, 96), np.uint8))
    cam = H.camera((10.5, 100.0, 10.5), 0.0, -10.0, 64, 40, W.DEFAULT_FLAGS)
    b.set_cameras([cam], [(W.DEFAULT_FLAGS & W.SPEC_MASK, W.DEFAULT_SPEC_HI)])
    assert b.view_counts(0) == (0, 0, A.VX_NO_CHUNK)
    b.close()
    r.close()
