"""vx_render_noalloc (the tile march on caller-held buffers, working memory in a fixed
device array) against vx_render, bit for bit: two views in a row, the second smaller,
two frames each with the Hi-Z words carried, and the first view against the oracle too.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.renderer import Renderer
from voxelcraft_menu_amd.world import build_view

pytestmark = pytest.mark.gpu

# The calls below run on the NULL stream; the context's copies run on its own non-blocking
# stream, which does not wait for NULL-stream work, so each call is followed by a device-wide
# wait before its outputs are read.
def _device_wait():
    assert C.CDLL("/opt/rocm/lib/libamdhip64.so").hipDeviceSynchronize() == 0


class _Dev:
    """Device buffers allocated through the context (at set-up, outside the calls tested)."""

    def __init__(self, r):
        self.L, self.c, self.bufs = A.march(), r.handle, []

    def alloc(self, nbytes):
        p = C.c_void_p()
        A.check(self.L.vx_dev_alloc(self.c, max(int(nbytes), 256), C.byref(p)), "vx_dev_alloc")
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        if a.nbytes:
            A.check(self.L.vx_dev_upload(self.c, p, A.ptr(a), a.nbytes), "vx_dev_upload")
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        A.check(self.L.vx_dev_download(self.c, A.ptr(out), p, out.nbytes), "vx_dev_download")
        return out

    def free(self):
        for p in self.bufs:
            self.L.vx_dev_free(self.c, p)
        self.bufs = []


def _noalloc_frames(dev, world, view, spec, hi, frames):
    pools = [np.ascontiguousarray(world.pool(i)) for i in range(3)]
    dv = A.DeviceView()
    dv.inners, dv.inners_bytes = dev.put(pools[0]).value, pools[0].nbytes
    dv.leaves, dv.leaves_bytes = dev.put(pools[1]).value, pools[1].nbytes
    dv.bricks, dv.bricks_bytes = dev.put(pools[2]).value, pools[2].nbytes
    dv.chunks, dv.n_chunks = dev.put(view.chunks).value, view.n_total
    dv.grid = dev.put(view.grid).value
    w, h = int(view.frame.res[0]), int(view.frame.res[1])
    nt = view.tiles[0] * view.tiles[1]
    hiz = dev.put(np.zeros(nt, np.uint32))
    vis = dev.alloc(w * h * 8)
    sh = dev.alloc(w * h * 2)
    out = []
    for _ in range(frames):
        A.check(dev.L.vx_render_noalloc(C.byref(view.frame), spec, hi, C.byref(dv), hiz, vis, sh, None),
                "vx_render_noalloc")
        _device_wait()
        out.append((dev.get(vis, np.zeros(w * h, np.uint64)), dev.get(sh, np.zeros(w * h, np.uint16)),
                    dev.get(hiz, np.zeros(nt, np.uint32))))
    return out, dv, hiz, vis, sh


def test_noalloc_matches_vx_render_on_two_views_the_second_smaller():
    world = W.make_world(W.WORKLOADS[0])
    cams = [W.default_camera(W.WORKLOADS[0], width=640, height=360),
            W.random_cameras(W.WORKLOADS[0], 1, seed=5, resolutions=[(320, 180)])[0]]
    r = Renderer(cams[0].width, cams[0].height)
    r.sync_world(world)
    dev = _Dev(r)
    try:
        for k, cam in enumerate(cams):
            if (cam.width, cam.height) != r.size:
                r.resize(cam.width, cam.height)
            v = build_view(world, cam)
            spec, hi = W.spec_words(cam.flags, world)
            print(f"noalloc view {k}: {cam.width}x{cam.height} pos={cam.pos}", flush=True)
            r.set_view(v)
            r.write_hiz(np.zeros(v.tiles[0] * v.tiles[1], np.uint32))
            want = []
            for _ in range(2):
                r.render(v.frame, spec, hi)
                r.sync()
                want.append((r.read_vis(), r.read_shadow(), r.read_hiz()))
            got, dv, *_ = _noalloc_frames(dev, world, v, spec, hi, 2)
            for f in range(2):
                for name, a, b in zip(("vis", "shadow", "hiz"), got[f], want[f]):
                    assert np.array_equal(a, b), f"view {k} frame {f}: {(a != b).sum()} {name} words differ"
            assert (got[1][0] != 0).sum() > 1000
            if k == 0:
                ref = H.oracle_frames(world, v, spec, hi, frames=2)
                for f in range(2):
                    assert np.array_equal(got[f][0], ref[f][0]) and np.array_equal(got[f][1], ref[f][1])
                    assert np.array_equal(got[f][2], ref[f][2])
    finally:
        dev.free()
        r.close()
