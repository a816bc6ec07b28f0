"""Shared test helpers: small worlds, chunk records, brute-force references."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from oracle import oracle as O
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.world import CameraParams, World, build_view

VOL = 64 ** 3


def dense_index(x, y, z):
    return x | (z << 6) | (y << 12)


def dense_from(fn) -> np.ndarray:
    """Dense chunk from fn(x, y, z) -> id, vectorised over the 64^3 grid."""
    y, z, x = np.meshgrid(np.arange(64), np.arange(64), np.arange(64), indexing="ij")
    out = np.asarray(fn(x, y, z), dtype=np.uint16)
    return np.broadcast_to(out, (64, 64, 64)).reshape(-1).copy()  # index y<<12 | z<<6 | x


def gpu_chunk(world: World, key) -> np.ndarray:
    """The 96-byte GpuChunk for a resident LOD-0 chunk (Renderer::build_chunk_list layout)."""
    info = world.chunk_info(key)
    ch = A.Chunk()
    o = [k * 64 for k in key]
    for i in range(3):
        ch.aabb_min[i] = o[i] + info.aabb_min[i]
        ch.aabb_max[i] = o[i] + info.aabb_max[i]
        ch.origin[i] = o[i]
    ch.voxel_size = 1.0
    ch.fade = 1.0
    for i in range(4):
        ch.root[i] = info.root[i]
    ch.attr_base = info.attr_base
    flags = 1 if info.attr_uniform else 0
    flags |= 2 if info.has_water else 0
    flags |= 4 if info.has_foliage else 0
    flags |= 8 if info.has_cutout else 0
    flags |= 16 if info.has_glass else 0
    flags |= 32 if info.has_emitter else 0
    ch.attr_flags = flags
    ch.light_base, ch.light_flags = 0xF000, 1
    ch.dry_mask[0] = info.dry_mask & 0xFFFFFFFF
    ch.dry_mask[1] = info.dry_mask >> 32
    return np.frombuffer(bytes(ch), dtype=np.uint8).copy()


def scene_for(world: World, chunks: np.ndarray, grid: np.ndarray | None = None) -> O.Scene:
    if grid is None:
        grid = np.full(A.VX_GRID_LEN, A.VX_NO_CHUNK, np.uint32)
    return O.Scene(world.pool(0), world.pool(1), world.pool(2), chunks.reshape(-1, 96), grid)


def frame_for(width=64, height=64, far=2600.0, grid_min=(0, 0, 0)) -> A.Frame:
    f = A.Frame()
    f.res[0], f.res[1] = width, height
    f.tiles[0], f.tiles[1] = (width + 7) // 8, (height + 7) // 8
    f.far_ = far
    f.shadow_dist = 220.0
    for i in range(3):
        f.grid_min[i] = grid_min[i]
    f.grid_dim[0], f.grid_dim[1], f.grid_dim[2] = 32, 8, 32
    return f


def march_chunk(scene: O.Scene, frame: A.Frame, spec: O.Spec, ci: int, ro, rd, t_limit=1e9,
                skip_water=False, skip_glass=False, sun_ray=False) -> O.Hit:
    h = O.Hit()
    fb = np.frombuffer(bytes(frame), np.uint8).copy()
    ro_a = np.asarray(ro, np.float32)
    rd_a = np.asarray(rd, np.float32)
    O.lib().vxo_march_chunk(C.byref(scene.pools), C.c_void_p(scene.chunks.ctypes.data), scene.n_chunks,
                            C.c_void_p(scene.grid.ctypes.data), C.c_void_p(fb.ctypes.data), C.byref(spec), ci,
                            C.c_void_p(ro_a.ctypes.data), C.c_void_p(rd_a.ctypes.data), t_limit,
                            int(skip_water), 0, int(skip_glass), int(sun_ray), C.byref(h))
    return h


def brute_dda(solid: np.ndarray, ro, rd, max_steps=400):
    """Amanatides-Woo unit-voxel DDA in float64 over a dense 64^3 occupancy grid
    (indexed [y, z, x]); the same first-hit rule as Player::pick (src/player.rs:334-392).
    Returns (voxel xyz, face normal id, t) of the first solid voxel, or None."""
    ro = np.asarray(ro, np.float64)
    rd = np.asarray(rd, np.float64)
    # Enter the [0,64)^3 box.
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / rd
        t0 = (0.0 - ro) * inv
        t1 = (64.0 - ro) * inv
    tmin = np.minimum(t0, t1)
    tmax = np.maximum(t0, t1)
    t_enter, t_exit = max(tmin.max(), 0.0), tmax.min()
    if t_enter >= t_exit:
        return None
    axis = int(np.argmax(tmin)) if tmin.max() > 0 else -1
    p = ro + rd * t_enter
    cell = np.floor(p).astype(np.int64)
    if axis >= 0:
        cell[axis] = 0 if rd[axis] > 0 else 63
    cell = np.clip(cell, 0, 63)
    step = np.where(rd > 0, 1, -1)
    nxt = np.where(rd > 0, cell + 1, cell).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        tmax_v = np.where(rd != 0, (nxt - ro) * inv, np.inf)
        tdelta = np.where(rd != 0, np.abs(inv), np.inf)
    t = t_enter
    for _ in range(max_steps):
        if np.any(cell < 0) or np.any(cell > 63):
            return None
        if solid[cell[1], cell[2], cell[0]]:
            if axis < 0:
                return tuple(int(c) for c in cell), None, t
            nid = axis * 2 + (0 if rd[axis] > 0 else 1)
            return tuple(int(c) for c in cell), nid, t
        axis = int(np.argmin(tmax_v))
        t = tmax_v[axis]
        cell[axis] += step[axis]
        tmax_v[axis] += tdelta[axis]
    return None


def camera(pos, yaw_deg, pitch_deg, w, h, flags, sea=128.0, time_of_day=0.30):
    return CameraParams(pos=pos, yaw=math.radians(yaw_deg), pitch=math.radians(pitch_deg), width=w, height=h,
                        flags=flags, sea_level=sea, time_of_day=time_of_day)


def oracle_frames(world, view, spec, spec_hi, frames=1, hiz=None):
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), view.chunks, view.grid)
    sp = O.spec_from(spec, spec_hi)
    hz = np.zeros(view.tiles[0] * view.tiles[1], np.uint32) if hiz is None else hiz
    out = []
    for _ in range(frames):
        vis, sh, hz, st = O.render(sc, view.frame_bytes, sp, hz)
        out.append((vis, sh, hz.copy(), st))
    return out


__all__ = ["build_view", "World"]
