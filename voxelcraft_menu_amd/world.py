"""Host world store and per-view setup (Python face of libvxhost, include/vxhost.h).

Mirrors the reference's CPU side of the path: ``World`` (src/voxel/world.rs),
``build_chunk_list`` (src/render/mod.rs:1447-1564), the GpuFrame packing of
``Renderer::render`` (src/render/mod.rs:1612-1710), ``scene::sun`` and
``scene::camera_basis`` (src/scene.rs:10-15, :334-339).
"""
from __future__ import annotations

import ctypes as C
import math
import threading
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
    # A key is a chunk position (x, y, z); lod > 0 names the chunk of 64^3 voxels of
    # 2^lod blocks at that position (ChunkKey, src/voxel/chunk.rs:6-60).
    def insert_dense(self, key, dense: np.ndarray, lod: int = 0):
        d = np.ascontiguousarray(dense, dtype=np.uint16).reshape(-1)
        if d.size != 64 ** 3:
            raise ValueError("a chunk is 64^3 ids")
        if lod:
            A.check(self._L.vxw_insert_dense_key(self._h, C.byref(chunk_key(key, lod)), A.ptr(d)), "insert_dense")
        else:
            A.check(self._L.vxw_insert_dense(self._h, *key, A.ptr(d)), "insert_dense")

    def insert_local_flat(self, key, record: bytes):
        """insert_local of a flat LocalTree record (build_local_flat, Renderer.build_local)."""
        buf = C.create_string_buffer(bytes(record), len(record))
        A.check(self._L.vxw_insert_local_flat(self._h, *key, buf, len(record)), "insert_local_flat")

    def remove_chunk(self, key, lod: int = 0) -> bool:
        return self._L.vxw_remove_chunk_key(self._h, C.byref(chunk_key(key, lod))) == 0

    def contains(self, key, lod: int = 0) -> bool:
        info = A.ChunkInfo()
        return self._L.vxw_chunk_info_key(self._h, C.byref(chunk_key(key, lod)), C.byref(info)) == 0

    def get_block(self, x, y, z) -> int:
        return int(self._L.vxw_get_block(self._h, x, y, z))

    def set_block(self, x, y, z, block_id) -> bool:
        return self._L.vxw_set_block(self._h, x, y, z, block_id) == 1

    def apply_edits(self, edits: np.ndarray) -> int:
        e = np.ascontiguousarray(edits, dtype=A.EDIT_DTYPE)
        return int(self._L.vxw_apply_edits(self._h, A.ptr(e), e.size))

    def extract_dense(self, key, lod: int = 0) -> np.ndarray:
        out = np.zeros(64 ** 3, dtype=np.uint16)
        A.check(self._L.vxw_extract_dense_key(self._h, C.byref(chunk_key(key, lod)), A.ptr(out)), "extract_dense")
        return out

    def chunk_info(self, key, lod: int = 0) -> A.ChunkInfo:
        info = A.ChunkInfo()
        A.check(self._L.vxw_chunk_info_key(self._h, C.byref(chunk_key(key, lod)), C.byref(info)), "chunk_info")
        return info

    def keys(self) -> np.ndarray:
        """LOD-0 chunk positions in insertion order, (n, 3) int32."""
        n = self._L.vxw_chunk_keys(self._h, None, 0)
        out = np.zeros((n, 3), dtype=np.int32)
        self._L.vxw_chunk_keys(self._h, A.ptr(out), n)
        return out

    def keys_all(self) -> np.ndarray:
        """Chunk keys of every LOD in insertion order (KEY_DTYPE records)."""
        n = self._L.vxw_chunk_keys_all(self._h, None, 0)
        out = np.zeros(n, dtype=A.KEY_DTYPE)
        if n:
            self._L.vxw_chunk_keys_all(self._h, A.ptr(out), n)
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

    def dense(self, key, lod: int = 0) -> np.ndarray:
        """The chunk at key; a coarse key gets a stand-in sampled one column per voxel."""
        out = np.zeros(64 ** 3, dtype=np.uint16)
        A.check(A.host().vxw_generate_dense_key(C.byref(self.to_c()), C.byref(chunk_key(key, lod)), A.ptr(out)),
                "generate_dense")
        return out


PROBE_AXES, PROBE_PER_AXIS = 6, 16  # probe::AXES, probe::PER_AXIS (src/probe.rs:6-12)


def probe_bake(gen: GenParams, origin, bounce_shadow: bool = True):
    """probe::bake of the chunk at `origin` with this generator's heights and surface
    blocks: (occl (6, 16, 16, 16), bounce (6, 16, 16, 16, 3)), indexed [axis, pz, py, px]."""
    o = np.asarray(origin, dtype=np.int32)
    occl = np.zeros((PROBE_AXES, 16, 16, 16), np.float32)
    bounce = np.zeros((PROBE_AXES, 16, 16, 16, 3), np.float32)
    A.check(A.host().vxw_probe_bake(C.byref(gen.to_c()), A.ptr(o), 1 if bounce_shadow else 0, A.ptr(occl),
                                    A.ptr(bounce)), "vxw_probe_bake")
    return occl, bounce


def probe_cube_from_tangents(tan_h) -> np.ndarray:
    t = np.ascontiguousarray(tan_h, dtype=np.float32)
    out = np.zeros(6, np.float32)
    A.host().vxw_probe_cube_from_tangents(A.ptr(t), A.ptr(out))
    return out


def probe_bounce_of_uniform_ground(albedo) -> np.ndarray:
    a = np.ascontiguousarray(albedo, dtype=np.float32)
    out = np.zeros((6, 3), np.float32)
    A.host().vxw_probe_bounce_of_uniform_ground(A.ptr(a), A.ptr(out))
    return out


def probe_uploadable(origin, cam) -> bool:
    o = np.asarray(origin, dtype=np.int32)
    c = np.asarray(cam, dtype=np.float32)
    return bool(A.host().vxw_probe_uploadable(A.ptr(o), A.ptr(c)))


def surface_block(gen: GenParams, x: int, z: int, h: int) -> int:
    return int(A.host().vxw_surface_block(C.byref(gen.to_c()), x, z, h))


def albedo(block_id: int) -> np.ndarray:
    return np.array([A.host().vxw_albedo(block_id, c) for c in range(3)], np.float32)


TINT_TEMP_FREQ, TINT_HUMID_FREQ = 0.00055, 0.0007  # biome::TEMP_FREQ / HUMID_FREQ (src/biome.rs:59-60)


def set_tint(frame: A.Frame, seed: int, strength: float = 1.0):
    """The frame's biome-tint fields as Renderer::render packs them (src/render/mod.rs:
    1686-1689): strength clamped to [0, 1], the world seed, the biome frequencies."""
    frame.tint_strength = min(max(float(strength), 0.0), 1.0)
    frame.tint_seed = int(seed)
    frame.tint_freq_t = TINT_TEMP_FREQ
    frame.tint_freq_h = TINT_HUMID_FREQ


class ShaftField:
    """shaft::ShaftField (src/shaft.rs): the 512 x 512 window of column heights, one texel
    per texel x texel blocks, centred on the camera, that the distant terrain shadows scan
    (vxw_shaft_*)."""

    DIM = 512
    SOFT = 4.0

    def __init__(self, texel: int = 4):
        self._L = A.host()
        self._s = self._L.vxw_shaft_new(int(texel))
        if not self._s:
            raise ValueError(f"shaft texel {texel} outside 1..64")

    def close(self):
        if self._s:
            self._L.vxw_shaft_free(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, gen: GenParams, cam_x: float, cam_z: float, canopy_lift: float = 0.0):
        A.check(self._L.vxw_shaft_update(self._s, C.byref(gen.to_c()), cam_x, cam_z, canopy_lift),
                "vxw_shaft_update")

    def heights(self) -> np.ndarray:
        p = self._L.vxw_shaft_heights(self._s)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(self.DIM, self.DIM)).copy()

    def origin_world(self):
        out = (C.c_float * 2)()
        self._L.vxw_shaft_origin_world(self._s, out)
        return float(out[0]), float(out[1])

    def texel(self) -> int:
        return int(self._L.vxw_shaft_texel(self._s))

    def dirty(self) -> bool:
        return bool(self._L.vxw_shaft_dirty(self._s))

    def envelope_reference(self, sun, tx: int, tz: int) -> float:
        s = np.asarray(sun, dtype=np.float32)
        return float(self._L.vxw_shaft_envelope(self._s, A.ptr(s), tx, tz))

    def set_frame(self, frame: A.Frame):
        """The frame's window fields, as Renderer::render writes them (src/render/mod.rs:
        1695-1697)."""
        o = self.origin_world()
        frame.shaft_origin[0], frame.shaft_origin[1] = o
        frame.shaft_texel = float(self.texel())
        frame.shaft_soft = self.SOFT


def chunk_key(key, lod: int = 0) -> A.Key:
    if not 0 <= lod <= A.VXW_MAX_LOD:
        raise ValueError(f"lod {lod} outside 0..{A.VXW_MAX_LOD}")
    return A.Key((C.c_int32 * 3)(*(int(v) for v in key)), lod)


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
    anim_time: float = 0.0                  # frame.time (config.rs anim_time)
    waves: tuple = (0.12, 32.0, 0.32, 0.0)  # wave amp, scale, speed, reflect slope (config.rs:318-321)

    def to_c(self) -> A.Camera:
        f, r, u = camera_basis(self.yaw, self.pitch)
        sd, dl = sun(self.time_of_day)
        return A.Camera((C.c_float * 3)(*self.pos), (C.c_float * 3)(*f), (C.c_float * 3)(*r), (C.c_float * 3)(*u),
                        math.radians(self.fov_deg), self.width, self.height, self.far,
                        (C.c_float * 3)(*sd), dl, self.sea_level, self.anim_time, (C.c_float * 2)(*self.jitter),
                        self.flags, self.frame_index, (C.c_float * 4)(*self.waves))


_tls = threading.local()


def _chunk_scratch() -> np.ndarray:
    """A per-thread buffer of VX_MAX_CHUNKS chunk records for the builders to write into;
    callers copy out the entries the builder returns."""
    buf = getattr(_tls, "chunks", None)
    if buf is None:
        buf = _tls.chunks = np.empty((A.VX_MAX_CHUNKS, 96), dtype=np.uint8)
    return buf


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
    chunks = _chunk_scratch()
    grid = np.empty(A.VX_GRID_LEN, dtype=np.uint32)  # ... and every grid cell
    v = A.View()
    A.check(L.vxw_build_view(world.handle, C.byref(c), A.ptr(chunks), A.ptr(grid), C.byref(v)), "build_view")
    fr = A.Frame()
    A.check(L.vxw_pack_frame(C.byref(c), C.byref(v), max_pairs, C.byref(fr)), "pack_frame")
    n = int(v.n_total)
    return ViewData(chunks[:n].copy() if n else np.zeros((1, 96), np.uint8), grid, fr,
                    int(v.n_visible), n, int(v.camera_chunk), cam.width, cam.height)


@dataclass
class LodConfig:
    """LodConfig (src/lod.rs:20-53) with its defaults."""
    max_lod: int = 4
    factor: float = 2.0
    view_distance: float = 2600.0
    fade_band: float = 1.25
    smootherstep: bool = False
    cover_depth: int = 2
    cross_fade: bool = True
    shadow_share: bool = True
    offscreen_shadows: bool = True

    def to_c(self) -> A.LodConfig:
        return A.LodConfig(self.max_lod, self.factor, self.view_distance, self.fade_band, int(self.smootherstep),
                           self.cover_depth, int(self.cross_fade), int(self.shadow_share),
                           int(self.offscreen_shadows))


class Planner:
    """lod::Planner (src/lod.rs:100-152): a camera's draw list over chunks of several LODs.

    ``plan`` returns (render, desired): RENDER_ITEM_DTYPE records (fade 1 solid, f a
    fading child, -f its parent) and the KEY_DTYPE keys the plan wants resident.
    Residency is read from ``world``, else from ``status(pos, lod) -> VXW_*``, else
    every key is solid. The planner keeps its split set between calls.
    """

    def __init__(self):
        self._L = A.host()
        self._h = self._L.vxw_planner_new()
        if not self._h:
            raise MemoryError("vxw_planner_new")

    def close(self):
        if self._h:
            self._L.vxw_planner_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def plan(self, cam, cfg: LodConfig | None = None, world: World | None = None, status=None):
        cfg = cfg or LodConfig()
        cam_a = (C.c_float * 3)(*cam)
        fn = A.RESIDENCY_FN()
        if world is None and status is not None:
            fn = A.RESIDENCY_FN(lambda _u, k: int(status(tuple(k.contents.pos), int(k.contents.lod))))
        A.check(self._L.vxw_planner_plan(self._h, cam_a, C.byref(cfg.to_c()), world.handle if world else None,
                                         fn, None), "planner_plan")
        n = self._L.vxw_planner_render(self._h, None, 0)
        render = np.zeros(n, dtype=A.RENDER_ITEM_DTYPE)
        if n:
            self._L.vxw_planner_render(self._h, A.ptr(render), n)
        m = self._L.vxw_planner_desired(self._h, None, 0)
        desired = np.zeros(m, dtype=A.KEY_DTYPE)
        if m:
            self._L.vxw_planner_desired(self._h, A.ptr(desired), m)
        return render, desired


def settle(world: World, gen: GenParams, planner: Planner, cam, cfg: LodConfig | None = None,
           max_rounds: int = 64):
    """Stream chunks until the plan at ``cam`` is resident: each round plans, then
    generates and inserts every wanted key (and its ancestors up to max_lod) that is not
    resident, coarse levels first (the request order of ChunkManager::update,
    src/stream.rs:252-313, without its worker threads, budget or retention). Returns the
    settled draw list."""
    cfg = cfg or LodConfig()
    for _ in range(max_rounds):
        render, desired = planner.plan(cam, cfg, world=world)
        missing = {}
        for k in desired:
            pos, lod = tuple(int(v) for v in k["pos"]), int(k["lod"])
            while True:
                if not world.contains(pos, lod) and (pos, lod) not in missing:
                    missing[(pos, lod)] = None
                if lod >= cfg.max_lod:
                    break
                pos, lod = tuple(v >> 1 for v in pos), lod + 1
        if not missing:
            return render
        for pos, lod in sorted(missing, key=lambda pk: -pk[1]):
            world.insert_dense(pos, gen.dense(pos, lod), lod)
    raise RuntimeError("the plan never settled")


def build_view_items(world: World, cam: CameraParams, items: np.ndarray, cfg: LodConfig | None = None,
                     max_pairs: int = 0) -> ViewData:
    """Renderer::build_chunk_list over a LOD draw list (RENDER_ITEM_DTYPE records) + the
    GpuFrame packing."""
    L = A.host()
    c = cam.to_c()
    it = np.ascontiguousarray(items, dtype=A.RENDER_ITEM_DTYPE)
    chunks = _chunk_scratch()
    grid = np.empty(A.VX_GRID_LEN, dtype=np.uint32)  # ... and every grid cell
    v = A.View()
    A.check(L.vxw_build_view_items(world.handle, C.byref(c), A.ptr(it) if it.size else None, it.size,
                                   C.byref((cfg or LodConfig()).to_c()), A.ptr(chunks), A.ptr(grid), C.byref(v)),
            "build_view_items")
    fr = A.Frame()
    A.check(L.vxw_pack_frame(C.byref(c), C.byref(v), max_pairs, C.byref(fr)), "pack_frame")
    n = int(v.n_total)
    return ViewData(chunks[:n].copy() if n else np.zeros((1, 96), np.uint8), grid, fr,
                    int(v.n_visible), n, int(v.camera_chunk), cam.width, cam.height)


class ViewSnapshot:
    """The LOD-0 chunk records view setup reads, copied from a world (vxw_snapshot_*):
    ``build_view`` gives what ``world.build_view`` gave when ``take`` ran, while the world
    itself may be under edit. Read-only between takes; safe to read from many threads."""

    def __init__(self, world: World | None = None):
        self._L = A.host()
        self._h = self._L.vxw_snapshot_new()
        if not self._h:
            raise MemoryError("vxw_snapshot_new")
        if world is not None:
            self.take(world)

    def close(self):
        if self._h:
            self._L.vxw_snapshot_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def take(self, world: World):
        A.check(self._L.vxw_snapshot_take(self._h, world.handle), "snapshot_take")

    def any_emitter_resident(self) -> bool:
        return self._L.vxw_snapshot_emitter_chunks(self._h) != 0

    def chunk_table(self) -> np.ndarray:
        """The (n, 96) GpuChunk records the device builds view chunk lists from."""
        n = self._L.vxw_snapshot_chunks(self._h, None, 0)
        out = np.zeros((max(n, 1), 96), np.uint8)
        if n:
            self._L.vxw_snapshot_chunks(self._h, A.ptr(out), n)
        return out[:n]

    def chunk_table_all(self):
        """Every resident LOD-0 chunk's (96-byte) record, the empty ones included, and its
        solid voxel count: the table a device world starts from."""
        n = self._L.vxw_snapshot_chunks_all(self._h, None, None, 0)
        out = np.zeros((max(n, 1), 96), np.uint8)
        solid = np.zeros(max(n, 1), np.uint32)
        if n:
            self._L.vxw_snapshot_chunks_all(self._h, A.ptr(out), A.ptr(solid), n)
        return out[:n], solid[:n]

    def build_view(self, cam: CameraParams, max_pairs: int = 0) -> ViewData:
        L = A.host()
        c = cam.to_c()
        chunks = _chunk_scratch()
        grid = np.empty(A.VX_GRID_LEN, dtype=np.uint32)  # ... and every grid cell
        v = A.View()
        A.check(L.vxw_build_view_snapshot(self._h, C.byref(c), A.ptr(chunks), A.ptr(grid), C.byref(v)),
                "build_view_snapshot")
        fr = A.Frame()
        A.check(L.vxw_pack_frame(C.byref(c), C.byref(v), max_pairs, C.byref(fr)), "pack_frame")
        n = int(v.n_total)
        return ViewData(chunks[:n].copy() if n else np.zeros((1, 96), np.uint8), grid, fr,
                        int(v.n_visible), n, int(v.camera_chunk), cam.width, cam.height)


def camera_frame(cam: CameraParams, max_pairs: int = 0):
    """(GpuFrame, frustum planes) of a camera whose chunk list the device builds."""
    L = A.host()
    c = cam.to_c()
    fr = A.Frame()
    A.check(L.vxw_pack_camera_frame(C.byref(c), max_pairs, C.byref(fr)), "pack_camera_frame")
    planes = np.zeros(24, np.float32)
    A.check(L.vxw_frustum_planes(C.byref(c), A.ptr(planes)), "frustum_planes")
    return fr, planes


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


def flat_layout(inners: np.ndarray, leaves: np.ndarray, bricks: np.ndarray, chunks: np.ndarray):
    """The pools as one contiguous, 128-byte aligned block per chunk (vxw_flat_layout):
    returns the block words (u32) and the (n, 96) chunk records with their root run pointer
    and attr_base moved into them. Pool arrays are the u32 words of world.pool(i)."""
    L = A.host()
    ins = np.ascontiguousarray(inners, dtype=np.uint32).reshape(-1)
    lfs = np.ascontiguousarray(leaves, dtype=np.uint32).reshape(-1)
    brs = np.ascontiguousarray(bricks, dtype=np.uint32).reshape(-1)
    ch = np.ascontiguousarray(chunks, dtype=np.uint8).reshape(-1, 96)
    n = ch.shape[0]
    out = np.zeros((max(n, 1), 96), np.uint8)
    words = C.c_size_t()
    args = (A.ptr(ins) if ins.size else None, ins.size // 4, A.ptr(lfs) if lfs.size else None, lfs.size // 2,
            A.ptr(brs) if brs.size else None, brs.size, A.ptr(ch) if n else None, n, A.ptr(out))
    A.check(L.vxw_flat_layout(*args, None, 0, C.byref(words)), "flat_layout")
    flat = np.zeros(max(words.value, 32), np.uint32)
    A.check(L.vxw_flat_layout(*args, A.ptr(flat), flat.size, C.byref(words)), "flat_layout")
    return flat[:max(words.value, 32)], out[:n]


def relocate_view(view: "ViewData", table: np.ndarray, moved: np.ndarray) -> "ViewData":
    """The view with each chunk record replaced by its moved record from `moved` (the
    flat_layout output for `table`), matched by origin and LOD."""
    key = lambda rec: bytes(rec[28:44])  # lod, origin
    index = {key(table[i]): i for i in range(table.shape[0])}
    ch = view.chunks.copy()
    for j in range(view.n_total):
        ch[j] = moved[index[key(view.chunks[j])]]
    return ViewData(ch, view.grid, view.frame, view.n_visible, view.n_total, view.camera_chunk, view.width,
                    view.height)


NEAR = 0.05  # src/render/mod.rs:140


def view_proj(frame: A.Frame) -> np.ndarray:
    """math::view_proj (src/math.rs:42-67) of a frame's camera: the 4x4 matrix (row, col)
    a later frame's prev_view_proj holds, in float64."""
    pos = np.array(frame.cam_pos, np.float64)
    fwd = np.array(frame.cam_fwd, np.float64)
    right = np.array(frame.cam_right, np.float64)
    up = np.array(frame.cam_up, np.float64)
    view = np.array([[right[0], right[1], right[2], -right @ pos],
                     [up[0], up[1], up[2], -up @ pos],
                     [-fwd[0], -fwd[1], -fwd[2], fwd @ pos],
                     [0.0, 0.0, 0.0, 1.0]])
    f = 1.0 / float(frame.tan_half_fov)
    far = float(frame.far_)
    proj = np.array([[f / float(frame.aspect), 0, 0, 0], [0, f, 0, 0],
                     [0, 0, far / (NEAR - far), NEAR * far / (NEAR - far)], [0, 0, -1.0, 0]])
    return proj @ view


def set_prev_view_proj(frame: A.Frame, m: np.ndarray):
    """Store a (row, col) matrix in frame.prev_view_proj, column major as the GpuFrame
    holds a Mat4."""
    for c in range(4):
        for r in range(4):
            frame.prev_view_proj[c * 4 + r] = float(m[r, c])
