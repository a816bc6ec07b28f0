"""GPU renderer handle: the reference ``Renderer`` surface for the march passes.

``Renderer`` mirrors src/render/mod.rs: ``sync_world`` (:1355-1396) uploads the three
pools with GpuMirror semantics, ``build_chunk_list`` (:1447-1564) builds and uploads the
per-view chunk list and grid, ``render`` (:1593-1840) runs tile_select / march / Hi-Z /
recover / march / primary_shadow, and the ``read_*`` calls read results back
(:1948-2112). Every call goes to libvxmarch.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi as A
from .world import CameraParams, ViewData, World, build_view


class Renderer:
    def __init__(self, width: int, height: int, path: int = A.PATH_TILES, max_pairs: int = 0, device: int = -1):
        self._L = A.march()
        self._c = C.c_void_p()
        desc = A.CtxDesc(width, height, max_pairs, path, device)
        A.check(self._L.vx_ctx_create(C.byref(desc), C.byref(self._c)), "vx_ctx_create")
        self.size = (width, height)
        self.path = path
        self.view: ViewData | None = None
        self._frame = None
        self._batches: list = []

    def close(self):
        for b in list(self._batches):
            b.close()
        if self._c:
            self._L.vx_ctx_destroy(self._c)
            self._c = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._c

    def set_tuning(self, occupancy: int = 4, variant: int = 0, ordered: int = 1):
        A.check(self._L.vx_set_tuning(self._c, C.byref(A.Tuning(occupancy, variant, ordered))), "vx_set_tuning")

    def build_local(self, dense: np.ndarray) -> list[bytes]:
        """build_local of n dense chunks ((n, 64^3) u16) on the device: n flat LocalTree
        records, each equal to world.build_local_flat of the same chunk."""
        d = np.ascontiguousarray(dense, dtype=np.uint16).reshape(-1, 64 ** 3)
        n = d.shape[0]
        out = np.zeros(n * A.LOCAL_TREE_MAX, dtype=np.uint8)
        A.check(self._L.vx_build_local(self._c, A.ptr(d), n, A.ptr(out), A.LOCAL_TREE_MAX), "vx_build_local")
        recs = []
        for i in range(n):
            slot = out[i * A.LOCAL_TREE_MAX:(i + 1) * A.LOCAL_TREE_MAX]
            nbytes = int(slot[60:64].view(np.uint32)[0])
            recs.append(slot[:nbytes].tobytes())
        return recs

    def tuning(self) -> dict:
        t = A.Tuning()
        A.check(self._L.vx_get_tuning(self._c, C.byref(t)), "vx_get_tuning")
        return {"occupancy": int(t.occupancy), "variant": int(t.variant), "ordered": int(t.ordered)}

    def resize(self, width: int, height: int):
        A.check(self._L.vx_resize(self._c, width, height), "vx_resize")
        self.size = (width, height)

    def sync_world(self, world: World):
        """Upload the pools' dirty ranges (or the whole pool when it outgrows the mirror)."""
        for pool in (A.POOL_INNERS, A.POOL_LEAVES, A.POOL_BRICKS):
            dirty = world.take_dirty(pool)
            ptr, nbytes = world.pool_view(pool)
            A.check(self._L.vx_pool_sync(self._c, pool, ptr, nbytes, A.ptr(dirty) if len(dirty) else None,
                                         len(dirty)), "vx_pool_sync")

    def upload_pools(self, inners: np.ndarray, leaves: np.ndarray, bricks: np.ndarray):
        """Full upload of caller-held pool arrays (u32 words)."""
        for pool, arr in ((A.POOL_INNERS, inners), (A.POOL_LEAVES, leaves), (A.POOL_BRICKS, bricks)):
            a = np.ascontiguousarray(arr, dtype=np.uint32)
            n_elem = a.nbytes // A.POOL_ELEM[pool]
            rng = np.array([[0, n_elem]], dtype=np.uint32)
            A.check(self._L.vx_pool_sync(self._c, pool, A.ptr(a) if a.size else None, a.nbytes,
                                         A.ptr(rng) if n_elem else None, 1 if n_elem else 0), "vx_pool_sync")

    def set_flat_pools(self, words: np.ndarray | None):
        """Read every pool from one block per chunk (world.flat_layout's words), or, with
        None, from the mirrors sync_world keeps. Chunk records must match the layout."""
        if words is None:
            A.check(self._L.vx_pool_set_flat(self._c, None, 0), "vx_pool_set_flat")
            return
        w = np.ascontiguousarray(words, dtype=np.uint32)
        A.check(self._L.vx_pool_set_flat(self._c, A.ptr(w), w.nbytes), "vx_pool_set_flat")

    def set_view(self, view: ViewData):
        A.check(self._L.vx_set_chunks(self._c, A.ptr(view.chunks), view.n_total, A.ptr(view.grid)), "vx_set_chunks")
        self.view = view

    def build_chunk_list(self, world: World, cam: CameraParams, max_pairs: int = 0) -> ViewData:
        v = build_view(world, cam, max_pairs)
        self.set_view(v)
        return v

    def render(self, frame: A.Frame, spec: int, spec_hi: int):
        A.check(self._L.vx_render(self._c, C.byref(frame), spec, spec_hi), "vx_render")
        self._frame = frame

    def sync(self):
        A.check(self._L.vx_sync(self._c), "vx_sync")

    def read_vis(self) -> np.ndarray:
        w, h = self.size
        out = np.zeros(w * h, dtype=np.uint64)
        A.check(self._L.vx_read_vis(self._c, A.ptr(out), out.size), "vx_read_vis")
        return out

    def read_shadow(self) -> np.ndarray:
        w, h = self.size
        out = np.zeros(w * h, dtype=np.uint16)
        A.check(self._L.vx_read_shadow(self._c, A.ptr(out), out.size), "vx_read_shadow")
        return out

    def read_hiz(self) -> np.ndarray:
        w, h = self.size
        out = np.zeros(((w + 7) // 8) * ((h + 7) // 8), dtype=np.uint32)
        A.check(self._L.vx_read_hiz(self._c, A.ptr(out), out.size), "vx_read_hiz")
        return out

    def write_hiz(self, hiz: np.ndarray):
        a = np.ascontiguousarray(hiz, dtype=np.uint32)
        A.check(self._L.vx_write_hiz(self._c, A.ptr(a), a.size), "vx_write_hiz")

    def read_pairs(self, n: int):
        """(pair words[0:n], counters[16]) of the last frame; pairs-path contexts only."""
        out = np.zeros(max(n, 1), dtype=np.uint32)
        cnt = np.zeros(16, dtype=np.uint32)
        A.check(self._L.vx_read_pairs(self._c, A.ptr(out), n, A.ptr(cnt)), "vx_read_pairs")
        return out[:n], cnt

    def read_decoded(self) -> np.ndarray:
        w, h = self.size
        out = np.zeros(w * h, dtype=A.HIT_DTYPE)
        A.check(self._L.vx_read_decoded(self._c, A.ptr(out), w * h), "vx_read_decoded")
        return out

    def set_resolve_assets(self, atlas: np.ndarray, size: int, layers: int, mips: int, block_faces: np.ndarray):
        """The atlas (RGBA8 sRGB bytes: mip level 0 of every layer, then level 1, ...) and
        the block face table (8 u32 words per block id) the resolve samples."""
        at = np.ascontiguousarray(atlas, dtype=np.uint8).reshape(-1)
        fc = np.ascontiguousarray(block_faces, dtype=np.uint32).reshape(-1)
        a = A.ResolveAssets(at.ctypes.data, size, layers, mips, fc.ctypes.data, fc.size)
        A.check(self._L.vx_resolve_set_assets(self._c, C.byref(a)), "vx_resolve_set_assets")

    def resolve(self):
        """The resolve (resolve.wgsl:1378-1535: water_sec, then lanes 0, 1 and 2) of the last
        rendered frame."""
        A.check(self._L.vx_resolve(self._c), "vx_resolve")

    def read_color(self) -> np.ndarray:
        """(H, W, 4) float32 colour of the last resolve."""
        w, h = self.size
        out = np.zeros((h, w, 4), dtype=np.float32)
        A.check(self._L.vx_read_color(self._c, A.ptr(out), out.size), "vx_read_color")
        return out

    def fill_probe(self, value: float | None = None, noise: bool = False):
        """make_probe's fill of the whole probe field: 1.0, a pinned value, or its noise."""
        v = C.c_float(value) if value is not None else None
        A.check(self._L.vx_probe_fill(self._c, C.byref(v) if v is not None else None, 1 if noise else 0),
                "vx_probe_fill")

    def write_probe(self, origin, occl: np.ndarray, bounce: np.ndarray):
        """Renderer::write_probe: a chunk's bake (world.probe_bake) into the probe field."""
        o = np.asarray(origin, dtype=np.int32)
        oc = np.ascontiguousarray(occl, dtype=np.float32)
        bo = np.ascontiguousarray(bounce, dtype=np.float32)
        assert oc.size == 6 * 4096 and bo.size == 6 * 4096 * 3
        A.check(self._L.vx_probe_write(self._c, A.ptr(o), A.ptr(oc), A.ptr(bo)), "vx_probe_write")

    def read_probe(self) -> np.ndarray:
        """(128, 768, 128, 4) f16 bit patterns of the probe field, [z, y, x, rgba]."""
        out = np.zeros((128, 768, 128, 4), np.uint16)
        A.check(self._L.vx_read_probe(self._c, A.ptr(out), out.size), "vx_read_probe")
        return out

    def upload_shaft(self, heights: np.ndarray):
        """The shaft field's 512 x 512 column heights (Renderer::upload_shaft)."""
        h = np.ascontiguousarray(heights, dtype=np.float32).reshape(-1)
        A.check(self._L.vx_shaft_upload(self._c, A.ptr(h), h.size), "vx_shaft_upload")

    def shaft_scan(self, frame: A.Frame):
        """shaft_scan's eight passes with the frame's sun_dir and shaft_texel; the distant
        terrain shadows of the next resolve read the result."""
        A.check(self._L.vx_shaft_scan(self._c, C.byref(frame)), "vx_shaft_scan")

    def read_shaft(self) -> np.ndarray:
        """(512, 512) f32: the last scan's result."""
        out = np.zeros((512, 512), dtype=np.float32)
        A.check(self._L.vx_read_shaft(self._c, A.ptr(out), out.size), "vx_read_shaft")
        return out

    def taa(self):
        """The TAA pass (taa.wgsl) over the last resolve, with the frame's taa_feedback and
        prev_view_proj; its output becomes the history and what present draws."""
        A.check(self._L.vx_taa(self._c), "vx_taa")

    def read_taa(self) -> np.ndarray:
        """(H, W, 4) f16 bit patterns (uint16) of the last TAA output."""
        w, h = self.size
        out = np.zeros((h, w, 4), dtype=np.uint16)
        A.check(self._L.vx_read_taa(self._c, A.ptr(out), out.size), "vx_read_taa")
        return out

    def write_taa_history(self, hist: np.ndarray):
        """Replace the history the next taa() reads ((H, W, 4) f16 bit patterns)."""
        h = np.ascontiguousarray(hist, dtype=np.uint16)
        A.check(self._L.vx_write_taa_history(self._c, A.ptr(h), h.size), "vx_write_taa_history")

    def present(self, encode: int = 1, tone: int = 0, grade: int = 0, strength: float = 1.0):
        """The present blit (blit.wgsl) of the last resolve into an RGBA8 image."""
        d = A.PresentDesc(encode, tone, grade, strength)
        A.check(self._L.vx_present(self._c, C.byref(d)), "vx_present")

    def read_image(self) -> np.ndarray:
        """(H, W, 4) uint8 image of the last present."""
        w, h = self.size
        out = np.zeros((h, w, 4), dtype=np.uint8)
        A.check(self._L.vx_read_image(self._c, A.ptr(out), out.size), "vx_read_image")
        return out

    def read_heat(self) -> np.ndarray:
        """Marched iterations per pixel of the last FLAG_HEATMAP frame."""
        w, h = self.size
        out = np.zeros(w * h, dtype=np.uint32)
        A.check(self._L.vx_read_heat(self._c, A.ptr(out), out.size), "vx_read_heat")
        return out

    def march_stats(self) -> dict:
        """Renderer::march_stats (src/render/mod.rs:1977-2045) of the last heatmap frame."""
        m = A.MarchStats()
        A.check(self._L.vx_march_stats_get(self._c, C.byref(m)), "vx_march_stats_get")
        return A.march_stats_dict(m)

    def pair_demand(self) -> dict:
        d = A.PairDemand()
        A.check(self._L.vx_pair_demand_get(self._c, C.byref(d)), "vx_pair_demand_get")
        return {k: int(getattr(d, k)) for k, _ in A.PairDemand._fields_}

    def timings(self) -> dict:
        t = A.Timings()
        A.check(self._L.vx_timings_get(self._c, C.byref(t)), "vx_timings_get")
        return {k: float(getattr(t, k)) for k, _ in A.Timings._fields_}

    def timings_reset(self):
        A.check(self._L.vx_timings_reset(self._c), "vx_timings_reset")

    def timings_sum(self):
        """(per-pass ms summed since timings_reset, frame count)."""
        t = A.Timings()
        n = C.c_uint32()
        A.check(self._L.vx_timings_sum(self._c, C.byref(t), C.byref(n)), "vx_timings_sum")
        return {k: float(getattr(t, k)) for k, _ in A.Timings._fields_}, int(n.value)


class Batch:
    """Many views of the renderer's world per submission (vx_batch_*).

    No counterpart in the reference, whose request loop renders one view per
    ``Renderer::render`` call (src/headless.rs:680-837). Views share the renderer's
    pools; each keeps its own Hi-Z, visibility and shadow buffers, and its output bits
    equal ``Renderer.render`` on the same frame, chunk list, grid and Hi-Z input.
    """

    def __init__(self, renderer: Renderer, max_views: int):
        self._L = renderer._L
        self._r = renderer
        self._b = C.c_void_p()
        A.check(self._L.vx_batch_create(renderer.handle, max_views, C.byref(self._b)), "vx_batch_create")
        renderer._batches.append(self)
        self.views: list[ViewData] = []
        self._keep = None

    def close(self):
        if self._b:
            self._L.vx_batch_destroy(self._b)
            self._b = C.c_void_p()
            if self in self._r._batches:
                self._r._batches.remove(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_views(self, views: list[ViewData], specs: list[tuple[int, int]]):
        """Upload the views (their frames, chunk lists and grids) and their pipeline keys."""
        descs = (A.ViewDesc * len(views))()
        keep = []
        for i, (v, (spec, spec_hi)) in enumerate(zip(views, specs)):
            ch = np.ascontiguousarray(v.chunks[: v.n_total])
            fr = A.Frame.from_buffer_copy(bytes(v.frame))
            keep += [ch, fr, v.grid]
            descs[i] = A.ViewDesc(C.addressof(fr), ch.ctypes.data if v.n_total else None, v.n_total,
                                  v.grid.ctypes.data, spec, spec_hi)
        A.check(self._L.vx_batch_set_views(self._b, descs, len(views)), "vx_batch_set_views")
        self.views = list(views)
        self._keep = keep

    def set_chunk_table(self, table: np.ndarray):
        """Upload the LOD-0 chunk records (ViewSnapshot.chunk_table) set_cameras builds from."""
        t = np.ascontiguousarray(table, dtype=np.uint8).reshape(-1, 96)
        A.check(self._L.vx_batch_set_chunk_table(self._b, A.ptr(t) if len(t) else None, len(t)),
                "vx_batch_set_chunk_table")

    def set_cameras(self, cams: list, specs: list[tuple[int, int]], moved: bool = False):
        """Views from cameras alone: their chunk lists and grids are built on the device
        from the chunk table, as build_view builds them on the host. moved: the installed
        views one frame later (vx_batch_move_cameras), keeping their Hi-Z words."""
        from .world import camera_frame

        descs = (A.CameraView * len(cams))()
        frames = []
        for i, (cam, (spec, spec_hi)) in enumerate(zip(cams, specs)):
            fr, planes = camera_frame(cam)
            frames.append(fr)
            descs[i] = A.CameraView(C.addressof(fr), spec, spec_hi, (C.c_float * 24)(*planes))
        fn = self._L.vx_batch_move_cameras if moved else self._L.vx_batch_set_cameras
        A.check(fn(self._b, descs, len(cams)), "vx_batch_move_cameras" if moved else "vx_batch_set_cameras")
        views = []
        for i, (cam, fr) in enumerate(zip(cams, frames)):
            nt, nv, cc = self.view_counts(i)
            fr.chunk_count, fr.camera_chunk = nv, cc
            views.append(ViewData(np.zeros((1, 96), np.uint8), np.zeros(0, np.uint32), fr, nv, nt, cc,
                                  cam.width, cam.height))
        self.views = views
        self._keep = frames

    def view_counts(self, i: int) -> tuple[int, int, int]:
        """(records, visible records, camera chunk) of view i's list."""
        nt, nv, cc = C.c_uint32(), C.c_uint32(), C.c_uint32()
        A.check(self._L.vx_batch_view_counts(self._b, i, C.byref(nt), C.byref(nv), C.byref(cc)), "vx_batch_view_counts")
        return int(nt.value), int(nv.value), int(cc.value)

    def read_chunks(self, i: int) -> tuple[np.ndarray, np.ndarray]:
        """View i's chunk list (n, 96) and grid as the batch holds them on the device."""
        nt = self.view_counts(i)[0]
        ch = np.zeros((max(nt, 1), 96), np.uint8)
        grid = np.zeros(A.VX_GRID_LEN, np.uint32)
        A.check(self._L.vx_batch_read_chunks(self._b, i, A.ptr(ch), nt, A.ptr(grid)), "vx_batch_read_chunks")
        return ch[:nt], grid

    def set_tuning(self, mode: int = A.BATCH_PERSISTENT, occupancy: int = 4, variant: int = 4, ordered: int = 1,
                   queue_chunk: int = 1, waves_per_simd: int | None = None, graph: int = 1):
        wps = occupancy if waves_per_simd is None else waves_per_simd
        t = A.BatchTuning(mode, occupancy, variant, ordered, queue_chunk, wps, graph)
        A.check(self._L.vx_batch_set_tuning(self._b, C.byref(t)), "vx_batch_set_tuning")

    def tuning(self) -> dict:
        t = A.BatchTuning()
        A.check(self._L.vx_batch_get_tuning(self._b, C.byref(t)), "vx_batch_get_tuning")
        return {k: int(getattr(t, k)) for k, _ in A.BatchTuning._fields_}

    def autotune(self, reps: int = 3) -> tuple[dict, float]:
        t = A.BatchTuning()
        ms = C.c_float()
        A.check(self._L.vx_batch_autotune(self._b, reps, C.byref(t), C.byref(ms)), "vx_batch_autotune")
        return {k: int(getattr(t, k)) for k, _ in A.BatchTuning._fields_}, float(ms.value)

    def render(self):
        A.check(self._L.vx_batch_render(self._b), "vx_batch_render")

    def sync(self):
        A.check(self._L.vx_batch_sync(self._b), "vx_batch_sync")

    def reset_hiz(self):
        A.check(self._L.vx_batch_reset_hiz(self._b), "vx_batch_reset_hiz")

    def _size(self, i):
        v = self.views[i]
        return int(v.frame.res[0]), int(v.frame.res[1])

    def read_vis(self, i: int) -> np.ndarray:
        w, h = self._size(i)
        out = np.zeros(w * h, dtype=np.uint64)
        A.check(self._L.vx_batch_read_vis(self._b, i, A.ptr(out), out.size), "vx_batch_read_vis")
        return out

    def read_shadow(self, i: int) -> np.ndarray:
        w, h = self._size(i)
        out = np.zeros(w * h, dtype=np.uint16)
        A.check(self._L.vx_batch_read_shadow(self._b, i, A.ptr(out), out.size), "vx_batch_read_shadow")
        return out

    def read_hiz(self, i: int) -> np.ndarray:
        w, h = self._size(i)
        out = np.zeros(((w + 7) // 8) * ((h + 7) // 8), dtype=np.uint32)
        A.check(self._L.vx_batch_read_hiz(self._b, i, A.ptr(out), out.size), "vx_batch_read_hiz")
        return out

    def set_secondary(self, on: bool = True):
        """Trace the secondary legs of water / glass hits with every frame."""
        A.check(self._L.vx_batch_set_secondary(self._b, 1 if on else 0), "vx_batch_set_secondary")

    def read_secondary(self, i: int) -> np.ndarray:
        """(W*H, 4) leg records of view i: [:, 0] refract / transmit, [:, 1] reflect, and
        where the glass transmit leg ends on water, that water's [:, 2] refract and [:, 3]
        reflect legs."""
        w, h = self._size(i)
        out = np.zeros((w * h, 4), dtype=A.SEC_DTYPE)
        A.check(self._L.vx_batch_read_secondary(self._b, i, A.ptr(out), out.size), "vx_batch_read_secondary")
        return out

    def secondary_count(self) -> dict:
        wl, gl, sr, ov = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint32()
        A.check(self._L.vx_batch_secondary_count(self._b, C.byref(wl), C.byref(gl), C.byref(sr), C.byref(ov)),
                "vx_batch_secondary_count")
        return {"water_legs": int(wl.value), "glass_legs": int(gl.value), "sun_rays": int(sr.value),
                "overflow": int(ov.value)}

    def demand(self) -> dict:
        d = A.PairDemand()
        A.check(self._L.vx_batch_demand_get(self._b, C.byref(d)), "vx_batch_demand_get")
        return {k: int(getattr(d, k)) for k, _ in A.PairDemand._fields_}

    def timings_reset(self):
        A.check(self._L.vx_batch_timings_reset(self._b), "vx_batch_timings_reset")

    def timings_sum(self) -> tuple[float, float, int]:
        """(submission ms, march-kernel ms, frames) summed since timings_reset."""
        ms = C.c_float()
        mm = C.c_float()
        n = C.c_uint32()
        A.check(self._L.vx_batch_timings_sum(self._b, C.byref(ms), C.byref(mm), C.byref(n)), "vx_batch_timings_sum")
        return float(ms.value), float(mm.value), int(n.value)


class Server:
    """Continuous-batching view server over the renderer's resident world (vx_serve_*).

    No counterpart in the reference (its request loop renders and waits one view at a
    time, src/headless.rs:419-497). ``submit`` queues a view; a native worker thread
    admits queued views onto free launch lanes by view slots and HBM budget and launches
    each lane's views as one batch. ``poll`` returns completions with their timestamps;
    ``release`` frees a completed request's outputs. Pool updates (``Renderer.sync_world``
    after edits) go between ``barrier`` and ``resume``.
    """

    def __init__(self, renderer: Renderer, lanes: int = 2, max_views: int = 64, hbm_budget: int = 4 << 30,
                 mode: int = A.BATCH_FLAT, occupancy: int = 4, variant: int = 4, queue_chunk: int = 4,
                 waves_per_simd: int | None = None):
        self._L = renderer._L
        self._r = renderer
        self._s = C.c_void_p()
        wps = occupancy if waves_per_simd is None else waves_per_simd
        t = A.BatchTuning(mode, occupancy, variant, 0, queue_chunk, wps, 0)
        d = A.ServeDesc(lanes, max_views, hbm_budget, t)
        A.check(self._L.vx_serve_create(renderer.handle, C.byref(d), C.byref(self._s)), "vx_serve_create")
        renderer._batches.append(self)
        self._res = (A.ServeResult * 256)()

    def close(self):
        if self._s:
            self._L.vx_serve_destroy(self._s)
            self._s = C.c_void_p()
            if self in self._r._batches:
                self._r._batches.remove(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self) -> float:
        return float(self._L.vx_serve_clock())

    def submit(self, rid: int, view: ViewData, spec: int, spec_hi: int, t_arrival: float = 0.0):
        ch = np.ascontiguousarray(view.chunks[: view.n_total])
        fr = A.Frame.from_buffer_copy(bytes(view.frame))
        d = A.ViewDesc(C.addressof(fr), ch.ctypes.data if view.n_total else None, view.n_total,
                       view.grid.ctypes.data, spec, spec_hi)
        A.check(self._L.vx_serve_submit(self._s, rid, C.byref(d), t_arrival), "vx_serve_submit")

    def set_chunk_table(self, table: np.ndarray):
        """The LOD-0 chunk records camera requests build their lists from; only while no
        lane runs and no camera request is pending (between barrier and resume)."""
        t = np.ascontiguousarray(table, dtype=np.uint8).reshape(-1, 96)
        A.check(self._L.vx_serve_set_chunk_table(self._s, A.ptr(t) if len(t) else None, len(t)),
                "vx_serve_set_chunk_table")

    def submit_camera(self, rid: int, cam, spec: int, spec_hi: int, t_arrival: float = 0.0):
        """A request given by its camera: the device builds its chunk list and grid."""
        from .world import camera_frame

        fr, planes = camera_frame(cam)
        d = A.CameraView(C.addressof(fr), spec, spec_hi, (C.c_float * 24)(*planes))
        A.check(self._L.vx_serve_submit_camera(self._s, rid, C.byref(d), t_arrival), "vx_serve_submit_camera")

    def poll(self, wait_ms: int = 0) -> list[dict]:
        n = C.c_uint32()
        A.check(self._L.vx_serve_poll(self._s, self._res, len(self._res), C.byref(n), wait_ms), "vx_serve_poll")
        return [{k: getattr(self._res[i], k) for k, _ in A.ServeResult._fields_} for i in range(n.value)]

    def read_vis(self, rid: int, width: int, height: int) -> np.ndarray:
        out = np.zeros(width * height, dtype=np.uint64)
        A.check(self._L.vx_serve_read_vis(self._s, rid, A.ptr(out), out.size), "vx_serve_read_vis")
        return out

    def read_shadow(self, rid: int, width: int, height: int) -> np.ndarray:
        out = np.zeros(width * height, dtype=np.uint16)
        A.check(self._L.vx_serve_read_shadow(self._s, rid, A.ptr(out), out.size), "vx_serve_read_shadow")
        return out

    def release(self, rid: int):
        A.check(self._L.vx_serve_release(self._s, rid), "vx_serve_release")

    def barrier(self):
        A.check(self._L.vx_serve_barrier(self._s), "vx_serve_barrier")

    def resume(self):
        A.check(self._L.vx_serve_resume(self._s), "vx_serve_resume")

    def stats(self) -> dict:
        st = A.ServeStats()
        A.check(self._L.vx_serve_stats_get(self._s, C.byref(st)), "vx_serve_stats_get")
        return {k: int(getattr(st, k)) for k, _ in A.ServeStats._fields_}

    def replay_start(self, arrivals, cams, specs, first_id: int = 0, t0: float = 0.0, keep: bool = False) -> float:
        """Starts a native replay: camera request i arrives at t0 + arrivals[i] (sorted,
        seconds) with pipeline key specs[i]; a native driver thread submits them on
        time and a native collector releases their outputs. Meanwhile submit nothing and
        poll nothing; barrier / pool sync / set_chunk_table / resume are allowed. t0 0:
        50 ms after the requests are prepared. keep: outputs stay until release().
        Returns t0."""
        from .world import camera_frame

        n = len(cams)
        reqs = (A.ReplayReq * n)()
        frames = []
        for i, (t, cam, (spec, hi)) in enumerate(zip(arrivals, cams, specs)):
            fr, planes = camera_frame(cam)
            frames.append(fr)
            reqs[i].t = float(t)
            reqs[i].view = A.CameraView(C.addressof(fr), spec, hi, (C.c_float * 24)(*planes))
        if not t0:
            t0 = self.clock() + 0.05
        A.check(self._L.vx_serve_replay_start(self._s, reqs, n, first_id, t0, 1 if keep else 0),
                "vx_serve_replay_start")
        self._replay_n = n  # the library copied the requests and frames
        return t0

    def replay_progress(self) -> tuple[int, int]:
        sub, done = C.c_uint32(), C.c_uint32()
        A.check(self._L.vx_serve_replay_progress(self._s, C.byref(sub), C.byref(done)), "vx_serve_replay_progress")
        return int(sub.value), int(done.value)

    def replay_join(self) -> list[dict]:
        """Waits for the replay; one result dict per request, in request order."""
        n = self._replay_n
        out = (A.ServeResult * max(n, 1))()
        A.check(self._L.vx_serve_replay_join(self._s, out, n), "vx_serve_replay_join")
        return [{k: getattr(out[i], k) for k, _ in A.ServeResult._fields_} for i in range(n)]


def device_name() -> str:
    buf = C.create_string_buffer(256)
    rc = A.march().vx_device_name(buf, 256)
    return buf.value.decode() if rc == 0 else "none"


class DeviceWorld:
    """A world kept on the GPU in the per-chunk block layout, edited there (vx_dworld_*).
    The renderer's context reads its pools from it while it lives; views use the chunk
    records of `chunks()`."""

    def __init__(self, renderer: "Renderer", world: World, capacity_factor: float = 8.0, min_words: int = 1 << 22):
        from .world import ViewSnapshot, flat_layout
        self._L = A.march()
        self._r = renderer
        snap = ViewSnapshot(world)
        table, solid = snap.chunk_table_all()
        snap.close()
        flat, moved = flat_layout(world.pool(0), world.pool(1), world.pool(2), table)
        # Edits that carve into full L1 cells expand their blocks (no run is shared), so a
        # batch can need more words than the chunks it rebuilds had.
        cap = int(max(flat.size * capacity_factor, flat.size + min_words) + 31) // 32 * 32
        self._w = C.c_void_p()
        A.check(self._L.vx_dworld_create(renderer._c, A.ptr(flat), flat.size, cap, A.ptr(moved), A.ptr(solid),
                                         moved.shape[0], C.byref(self._w)), "vx_dworld_create")
        renderer._batches.append(self)  # closed before the context it reads

    def close(self):
        if self._w:
            self._L.vx_dworld_destroy(self._w)
            self._w = None
            if self in self._r._batches:
                self._r._batches.remove(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def edit(self, edits: np.ndarray):
        """Apply a batch of 16-byte edit records (x, y, z: i32; id: u16) in order."""
        e = np.ascontiguousarray(edits)
        A.check(self._L.vx_dworld_edit(self._w, A.ptr(e) if len(e) else None, len(e)), "vx_dworld_edit")

    def compact(self):
        """Rebuild every non-empty chunk into fresh, packed pools (vx_dworld_compact)."""
        A.check(self._L.vx_dworld_compact(self._w), "vx_dworld_compact")

    def chunks(self) -> np.ndarray:
        """(n, 96) records of the non-empty chunks, in table order."""
        n = C.c_uint32()
        self._L.vx_dworld_chunks(self._w, None, 0, C.byref(n))
        out = np.zeros((max(n.value, 1), 96), np.uint8)
        A.check(self._L.vx_dworld_chunks(self._w, A.ptr(out), n.value, C.byref(n)), "vx_dworld_chunks")
        return out[:n.value]

    def extract(self, key) -> np.ndarray:
        out = np.zeros(64 ** 3, np.uint16)
        A.check(self._L.vx_dworld_extract(self._w, int(key[0]), int(key[1]), int(key[2]), A.ptr(out)),
                "vx_dworld_extract")
        return out

    def stats(self) -> dict:
        st = A.DWorldStats()
        A.check(self._L.vx_dworld_stats_get(self._w, C.byref(st)), "vx_dworld_stats_get")
        return {k: getattr(st, k) for k, _ in A.DWorldStats._fields_}
