"""Host world store and per-view setup (Python face of libvxhost, include/vxhost.h).

Mirrors the reference's CPU side of the path: ``World`` (src/voxel/world.rs),
``build_chunk_list`` (src/render/mod.rs:1447-1564), the GpuFrame packing of
``Renderer::render`` (src/render/mod.rs:1612-1710), ``scene::sun`` and
``scene::camera_basis`` (src/scene.rs:10-15, :334-339).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A

AIR, STONE, DIRT, GRASS, SAND, GRAVEL, LOG, LEAVES, PLANKS, COBBLE, GLASS, GLOWSTONE, BEDROCK, WATER = range(14)
PINE_LEAVES = 16


class World:
    """Sparse voxel tree world with hash-consed run pools (src/voxel/world.rs:16-638)."""

    def __init__(self):
        self._L = A.host()
        self._h = self._L.vxw_world_new()
        if not self._h:
            raise MemoryError("vxw_world_new")

    def close(self):
        if self._h:
            self._L.vxw_world_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ---- store ----
    def insert_dense(self, key, dense: np.ndarray):
        d = np.ascontiguousarray(dense, dtype=np.uint16).reshape(-1)
        if d.size != 64 ** 3:
            raise ValueError("a chunk is 64^3 ids")
        A.check(self._L.vxw_insert_dense(self._h, *key, A.ptr(d)), "insert_dense")

    def insert_local_flat(self, key, record: bytes):
        """insert_local of a flat LocalTree record (build_local_flat, Renderer.build_local)."""
        buf = C.create_string_buffer(bytes(record), len(record))
        A.check(self._L.vxw_insert_local_flat(self._h, *key, buf, len(record)), "insert_local_flat")

    def remove_chunk(self, key) -> bool:
        return self._L.vxw_remove_chunk(self._h, *key) == 0

    def get_block(self, x, y, z) -> int:
        return int(self._L.vxw_get_block(self._h, x, y, z))

    def set_block(self, x, y, z, block_id) -> bool:
        return self._L.vxw_set_block(self._h, x, y, z, block_id) == 1

    def apply_edits(self, edits: np.ndarray) -> int:
        e = np.ascontiguousarray(edits, dtype=A.EDIT_DTYPE)
        return int(self._L.vxw_apply_edits(self._h, A.ptr(e), e.size))

    def extract_dense(self, key) -> np.ndarray:
        out = np.zeros(64 ** 3, dtype=np.uint16)
        A.check(self._L.vxw_extract_dense(self._h, *key, A.ptr(out)), "extract_dense")
        return out

    def chunk_info(self, key) -> A.ChunkInfo:
        info = A.ChunkInfo()
        A.check(self._L.vxw_chunk_info_get(self._h, *key, C.byref(info)), "chunk_info")
        return info

    def keys(self) -> np.ndarray:
        n = self._L.vxw_chunk_keys(self._h, None, 0)
        out = np.zeros((n, 3), dtype=np.int32)
        self._L.vxw_chunk_keys(self._h, A.ptr(out), n)
        return out

    def __len__(self):
        return int(self._L.vxw_chunk_count(self._h))

    def any_emitter_resident(self) -> bool:
        return self._L.vxw_emitter_chunks(self._h) != 0

    # ---- pools ----
    def pool(self, pool: int) -> np.ndarray:
        """Copy of a pool as u32 words."""
        nbytes = self._L.vxw_pool_bytes(self._h, pool)
        out = np.zeros(nbytes // 4, dtype=np.uint32)
        if nbytes:
            C.memmove(out.ctypes.data, self._L.vxw_pool_data(self._h, pool), nbytes)
        return out

    def pool_view(self, pool: int):
        """(pointer, bytes) of the live pool; valid until the next mutation."""
        return self._L.vxw_pool_data(self._h, pool), int(self._L.vxw_pool_bytes(self._h, pool))

    def take_dirty(self, pool: int) -> np.ndarray:
        n = self._L.vxw_take_dirty(self._h, pool, None, 0)
        out = np.zeros((n, 2), dtype=np.uint32)
        if n:
            self._L.vxw_take_dirty(self._h, pool, A.ptr(out), n)
        return out

    def pool_stats(self, pool: int) -> dict:
        s = A.PoolStats()
        A.check(self._L.vxw_pool_stats_get(self._h, pool, C.byref(s)), "pool_stats")
        return {"live_elems": s.live_elems, "referenced_elems": s.referenced_elems,
                "live_runs": s.live_runs, "len": s.len}

    # ---- generation ----
    def generate(self, params: "GenParams"):
        A.check(self._L.vxw_generate(self._h, C.byref(params.to_c())), "generate")


@dataclass
class GenParams:
    seed: int = 1
    sea_level: int = 128
    chunk_min: tuple = (0, 0, 0)
    chunk_max: tuple = (1, 1, 1)
    cave_fraction: float = 0.0
    tree_density: float = 0.004
    glass_fraction: float = 0.0
    leaf_extra: float = 0.0
    threads: int = 0

    def to_c(self) -> A.GenParams:
        return A.GenParams(self.seed, self.sea_level, (C.c_int32 * 3)(*self.chunk_min),
                           (C.c_int32 * 3)(*self.chunk_max), self.cave_fraction, self.tree_density,
                           self.glass_fraction, self.leaf_extra, self.threads)

    def height(self, x: int, z: int) -> int:
        return int(A.host().vxw_height(C.byref(self.to_c()), x, z))

    def dense(self, key) -> np.ndarray:
        out = np.zeros(64 ** 3, dtype=np.uint16)
        A.check(A.host().vxw_generate_dense(C.byref(self.to_c()), *key, A.ptr(out)), "generate_dense")
        return out


def camera_basis(yaw: float, pitch: float):
    """Player::look_dir + scene::camera_basis (src/player.rs:122-126, src/scene.rs:334-339)."""
    f, r, u = (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 3)()
    A.host().vxw_camera_basis(yaw, pitch, f, r, u)
    return tuple(f), tuple(r), tuple(u)


def sun(time_of_day: float):
    """scene::sun (src/scene.rs:10-15): (direction, daylight)."""
    d = (C.c_float * 3)()
    dl = C.c_float()
    A.host().vxw_sun(time_of_day, d, C.byref(dl))
    return tuple(d), float(dl.value)


def pairs_for(tiles_x: int, tiles_y: int) -> int:
    return int(A.host().vxw_pairs_for(tiles_x, tiles_y))


@dataclass
class CameraParams:
    pos: tuple
    yaw: float
    pitch: float
    width: int
    height: int
    fov_deg: float = 70.0
    far: float = 2600.0
    time_of_day: float = 0.30
    sea_level: float = 128.0
    flags: int = 0
    jitter: tuple = (0.0, 0.0)
    frame_index: int = 0

    def to_c(self) -> A.Camera:
        f, r, u = camera_basis(self.yaw, self.pitch)
        sd, dl = sun(self.time_of_day)
        return A.Camera((C.c_float * 3)(*self.pos), (C.c_float * 3)(*f), (C.c_float * 3)(*r), (C.c_float * 3)(*u),
                        math.radians(self.fov_deg), self.width, self.height, self.far,
                        (C.c_float * 3)(*sd), dl, self.sea_level, 0.0, (C.c_float * 2)(*self.jitter),
                        self.flags, self.frame_index)


@dataclass
class ViewData:
    """Everything one view uploads: chunk list, grid and frame record."""
    chunks: np.ndarray          # (n_total, 96) uint8
    grid: np.ndarray            # (8192,) uint32
    frame: A.Frame
    n_visible: int
    n_total: int
    camera_chunk: int
    width: int = field(default=0)
    height: int = field(default=0)

    @property
    def frame_bytes(self) -> bytes:
        return bytes(self.frame)

    @property
    def tiles(self):
        return int(self.frame.tiles[0]), int(self.frame.tiles[1])


def build_view(world: World, cam: CameraParams, max_pairs: int = 0) -> ViewData:
    """Renderer::build_chunk_list + the GpuFrame packing of Renderer::render."""
    L = A.host()
    c = cam.to_c()
    chunks = np.zeros((A.VX_MAX_CHUNKS, 96), dtype=np.uint8)
    grid = np.zeros(A.VX_GRID_LEN, dtype=np.uint32)
    v = A.View()
    A.check(L.vxw_build_view(world.handle, C.byref(c), A.ptr(chunks), A.ptr(grid), C.byref(v)), "build_view")
    fr = A.Frame()
    A.check(L.vxw_pack_frame(C.byref(c), C.byref(v), max_pairs, C.byref(fr)), "pack_frame")
    n = int(v.n_total)
    return ViewData(chunks[:max(n, 1)].copy() if n else np.zeros((1, 96), np.uint8), grid, fr,
                    int(v.n_visible), n, int(v.camera_chunk), cam.width, cam.height)


def build_local_flat(dense: np.ndarray) -> bytes:
    """build_local (src/voxel/tree.rs:89-204) of one dense chunk on the host, as the flat
    vx_local_tree record the device build writes."""
    L = A.host()
    d = np.ascontiguousarray(dense, dtype=np.uint16).reshape(-1)
    if d.size != 64 ** 3:
        raise ValueError("a chunk is 64^3 ids")
    n = C.c_size_t()
    A.check(L.vxw_build_local_flat(A.ptr(d), None, 0, C.byref(n)), "build_local_flat")
    out = C.create_string_buffer(n.value)
    A.check(L.vxw_build_local_flat(A.ptr(d), out, n.value, C.byref(n)), "build_local_flat")
    return out.raw[:n.value]
