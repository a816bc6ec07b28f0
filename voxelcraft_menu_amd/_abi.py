"""ctypes view of the two C ABIs (include/vxhost.h, include/vxmarch.h).

The shared objects are built in-tree by ``voxelcraft_menu_amd.build`` and loaded
from this directory. Loading fails loudly: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "libvxhost.so")
MARCH_LIB = os.path.join(HERE, "libvxmarch.so")
CALIB_LIB = os.path.join(HERE, "libvxcalib.so")

# ---- constants (include/vxmarch.h; src/render/mod.rs:20-37, :144-270) ----
VX_OK = 0
VX_TILE = 8
VX_MAX_CHUNKS = 8191
VX_CHUNK_TABLE_MAX = 8192
VX_MAX_PAIRS = 1 << 20
VX_MAX_TILES = 1 << 19
VX_GRID = (32, 8, 32)
VX_GRID_LEN = 32 * 8 * 32
VX_NO_CHUNK = 0xFFFFFFFF
POOL_INNERS, POOL_LEAVES, POOL_BRICKS = 0, 1, 2
POOL_ELEM = {POOL_INNERS: 16, POOL_LEAVES: 8, POOL_BRICKS: 4}
PATH_TILES, PATH_PAIRS = 0, 1

FLAG_HIZ = 1
FLAG_HEATMAP = 2
FLAG_SHADOWS = 4
FLAG_AO = 8
FLAG_WATER_REFLECT = 16
FLAG_WATER_REFRACT = 32
FLAG_UNDERWATER = 64
FLAG_TINT = 128
FLAG_FOLIAGE = 256
FLAG_WAVES = 512
FLAG_DISTANT_SHADOWS = 32768
FLAG_LEAF_CUTOUT = 65536
FLAG_GLASS = 131072
FLAG_WATER_SHADOW_CUT = 262144
FLAG_LEAF_THIN = 524288
FLAG_WATER_FAR = 2097152
FLAG_WATER_DARK = 4194304
FLAG_SNELL = 8388608
FLAG_HI_SKY_SPECULAR = 1
FLAG_HI_FULL_MARCH = 2
FLAG_HI_EMITTER_WORLD = 4
FLAG_HI_GLASS_REFLECT = 128
FLAG_HI_WATER_LOOK = 4096
FLAG_HI_SNELL_BEND = 256
FLAG_HI_WIND_SWAY = 32768
FLAG_HI_SHADOW_PASS = 524288
FLAG_HI_LEGACY_WATER = 2097152

ERRORS = {0: "ok", -1: "no device", -2: "out of memory", -3: "bad argument", -4: "overflow",
          -5: "unsupported", -6: "hip error", -7: "missing", -8: "busy"}


class Frame(C.Structure):
    """GpuFrame, 352 bytes (src/render/mod.rs:398-466)."""
    _fields_ = [
        ("cam_pos", C.c_float * 3), ("time", C.c_float),
        ("cam_fwd", C.c_float * 3), ("tan_half_fov", C.c_float),
        ("cam_right", C.c_float * 3), ("aspect", C.c_float),
        ("cam_up", C.c_float * 3), ("far_", C.c_float),
        ("sun_dir", C.c_float * 3), ("daylight", C.c_float),
        ("res", C.c_uint32 * 2), ("tiles", C.c_uint32 * 2),
        ("chunk_count", C.c_uint32), ("camera_chunk", C.c_uint32),
        ("flags", C.c_uint32), ("max_pairs", C.c_uint32),
        ("fog_density", C.c_float), ("fog_falloff", C.c_float), ("shadow_dist", C.c_float),
        ("ambient", C.c_float), ("fog_scatter", C.c_float), ("fog_g", C.c_float),
        ("fog_height", C.c_float), ("mip_bias", C.c_float),
        ("grid_min", C.c_int32 * 3), ("sea_level", C.c_float),
        ("grid_dim", C.c_uint32 * 3), ("water_absorb", C.c_float),
        ("jitter", C.c_float * 2), ("taa_feedback", C.c_float),
        ("frame_index", C.c_uint32),
        ("cloud_cover", C.c_float), ("cloud_height", C.c_float), ("cloud_scale", C.c_float),
        ("cloud_speed", C.c_float), ("godray_strength", C.c_float), ("godray_steps", C.c_uint32),
        ("godray_dist", C.c_float), ("cloud_shadow", C.c_float), ("tint_strength", C.c_float),
        ("tint_seed", C.c_int32), ("tint_freq_t", C.c_float), ("tint_freq_h", C.c_float),
        ("wave_amp", C.c_float), ("wave_scale", C.c_float), ("wave_speed", C.c_float),
        ("wave_reflect_slope", C.c_float),
        ("shaft_origin", C.c_float * 2), ("shaft_texel", C.c_float), ("shaft_soft", C.c_float),
        ("cloud_patch", C.c_float), ("cloud_relief", C.c_float), ("haze_warm", C.c_float),
        ("zenith_deep", C.c_float),
        ("prev_view_proj", C.c_float * 16),
    ]


class Chunk(C.Structure):
    """GpuChunk, 96 bytes (src/render/mod.rs:468-489)."""
    _fields_ = [
        ("aabb_min", C.c_float * 3), ("voxel_size", C.c_float),
        ("aabb_max", C.c_float * 3), ("lod", C.c_uint32),
        ("origin", C.c_float * 3), ("fade", C.c_float),
        ("root", C.c_uint32 * 4),
        ("attr_base", C.c_uint32), ("attr_flags", C.c_uint32),
        ("light_base", C.c_uint32), ("light_flags", C.c_uint32),
        ("dry_mask", C.c_uint32 * 2), ("reserved", C.c_uint32 * 2),
    ]


assert C.sizeof(Frame) == 352 and Frame.prev_view_proj.offset == 288 and Frame.grid_min.offset == 144
assert C.sizeof(Chunk) == 96 and Chunk.root.offset == 48 and Chunk.dry_mask.offset == 80


class Range(C.Structure):
    _fields_ = [("off", C.c_uint32), ("len", C.c_uint32)]


class Edit(C.Structure):
    """Journal edit record, 16 bytes (src/journal.rs:116-122)."""
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32), ("id", C.c_uint16), ("pad", C.c_uint16)]


assert C.sizeof(Edit) == 16

EDIT_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("id", "<u2"), ("pad", "<u2")])


LOCAL_TREE_MAX = 64 + 64 * 24 + 4096 * 8 + 4096 * 4 + 4096 * 49 * 4  # VX_LOCAL_TREE_MAX


class LocalTree(C.Structure):  # vx_local_tree (flat LocalTree record header)
    _fields_ = [("root_mask", C.c_uint64), ("dry_mask", C.c_uint64), ("n_l1", C.c_uint32), ("n_leaves", C.c_uint32),
                ("n_attrs", C.c_uint32), ("n_palette", C.c_uint32), ("uniform", C.c_uint32), ("flags", C.c_uint32),
                ("emitters", C.c_uint32), ("solid_count", C.c_uint32), ("leaf_count", C.c_uint32),
                ("aabb_min", C.c_uint8 * 3), ("aabb_max", C.c_uint8 * 3), ("pad", C.c_uint8 * 2),
                ("bytes", C.c_uint32)]


class Player(C.Structure):  # vxw_player (journal PlayerState, src/journal.rs:36-45)
    _fields_ = [("pos", C.c_float * 3), ("yaw", C.c_float), ("pitch", C.c_float), ("hotbar", C.c_uint32),
                ("fly", C.c_uint32), ("pad", C.c_uint32)]


class JournalInfo(C.Structure):  # vxw_journal_info
    _fields_ = [("version", C.c_uint32), ("flags", C.c_uint32), ("seed", C.c_int32), ("sea_level", C.c_int32),
                ("count", C.c_uint32), ("recovered", C.c_uint32), ("torn_bytes", C.c_uint32),
                ("has_player", C.c_uint32), ("has_bar", C.c_uint32), ("player", Player), ("bar", C.c_uint16 * 10)]


class ChunkInfo(C.Structure):
    _fields_ = [
        ("key", C.c_int32 * 3), ("root", C.c_uint32 * 4), ("attr_uniform", C.c_uint32),
        ("attr_base", C.c_uint32), ("attr_len", C.c_uint32),
        ("aabb_min", C.c_int32 * 3), ("aabb_max", C.c_int32 * 3),
        ("solid_count", C.c_uint32), ("leaf_count", C.c_uint32),
        ("has_water", C.c_uint32), ("has_foliage", C.c_uint32), ("has_cutout", C.c_uint32),
        ("has_glass", C.c_uint32), ("has_emitter", C.c_uint32), ("emitters", C.c_uint32),
        ("dry_mask", C.c_uint64), ("lod", C.c_uint32), ("pad", C.c_uint32),
    ]


class Key(C.Structure):  # vxw_key: ChunkKey {pos, lod}
    _fields_ = [("pos", C.c_int32 * 3), ("lod", C.c_uint32)]


class RenderItem(C.Structure):  # vxw_render_item
    _fields_ = [("key", Key), ("fade", C.c_float)]


RENDER_ITEM_DTYPE = np.dtype([("pos", "<i4", 3), ("lod", "<u4"), ("fade", "<f4")])
KEY_DTYPE = np.dtype([("pos", "<i4", 3), ("lod", "<u4")])


class LodConfig(C.Structure):  # vxw_lod_config
    _fields_ = [("max_lod", C.c_uint32), ("factor", C.c_float), ("view_distance", C.c_float),
                ("fade_band", C.c_float), ("fade_schedule", C.c_uint32), ("cover_depth", C.c_uint32),
                ("cross_fade", C.c_uint32), ("shadow_share", C.c_uint32), ("offscreen_shadows", C.c_uint32)]


RESIDENCY_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(Key))
VXW_ABSENT, VXW_EMPTY, VXW_SOLID = 0, 1, 2
VXW_MAX_LOD = 12


class DWorldStats(C.Structure):  # vx_dworld_stats
    _fields_ = [("edits", C.c_uint64), ("effective_edits", C.c_uint64), ("used_bytes", C.c_uint64),
                ("capacity_bytes", C.c_uint64), ("chunks_rebuilt", C.c_uint32), ("chunks", C.c_uint32),
                ("last_ms", C.c_float), ("compactions", C.c_uint32)]


class PoolMirror(C.Structure):  # vx_pool_mirror
    _fields_ = [("cap", C.c_uint64), ("len", C.c_uint64), ("uploaded_bytes", C.c_uint64),
                ("generation", C.c_uint32), ("pad", C.c_uint32)]


class PoolStats(C.Structure):
    _fields_ = [("live_elems", C.c_uint64), ("referenced_elems", C.c_uint64),
                ("live_runs", C.c_uint64), ("len", C.c_uint64)]


class GenParams(C.Structure):
    _fields_ = [
        ("seed", C.c_int32), ("sea_level", C.c_int32),
        ("chunk_min", C.c_int32 * 3), ("chunk_max", C.c_int32 * 3),
        ("cave_fraction", C.c_float), ("tree_density", C.c_float),
        ("glass_fraction", C.c_float), ("leaf_extra", C.c_float), ("threads", C.c_int32),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3), ("fwd", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3),
        ("fov_y", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32), ("far_", C.c_float),
        ("sun_dir", C.c_float * 3), ("daylight", C.c_float), ("sea_level", C.c_float), ("time", C.c_float),
        ("jitter", C.c_float * 2), ("flags", C.c_uint32), ("frame_index", C.c_uint32),
        ("waves", C.c_float * 4),
    ]


class View(C.Structure):
    _fields_ = [("n_visible", C.c_uint32), ("n_total", C.c_uint32), ("camera_chunk", C.c_uint32),
                ("grid_min", C.c_int32 * 3)]


class CtxDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_pairs", C.c_uint32),
                ("path", C.c_uint32), ("device", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("occupancy", C.c_uint32), ("variant", C.c_uint32), ("ordered", C.c_uint32)]


BATCH_PERSISTENT, BATCH_FLAT, BATCH_PLAIN = 0, 1, 2


class ViewDesc(C.Structure):
    """vx_view: one view of a batch (pointers into caller-owned arrays)."""
    _fields_ = [("frame", C.c_void_p), ("chunks", C.c_void_p), ("n_total", C.c_uint32),
                ("grid", C.c_void_p), ("spec", C.c_uint32), ("spec_hi", C.c_uint32)]


class CameraView(C.Structure):
    """vx_camera_view: a batch view given by its camera (chunk list built on the device)."""
    _fields_ = [("frame", C.c_void_p), ("spec", C.c_uint32), ("spec_hi", C.c_uint32), ("planes", C.c_float * 24)]


class ResolveAssets(C.Structure):
    """vx_resolve_assets: the caller's atlas (RGBA8 sRGB, mip level 0 of every layer, then
    level 1, ...) and block face table (8 words per block id)."""
    _fields_ = [("atlas", C.c_void_p), ("size", C.c_uint32), ("layers", C.c_uint32), ("mips", C.c_uint32),
                ("block_faces", C.c_void_p), ("n_faces", C.c_uint32)]


class PresentDesc(C.Structure):
    """vx_present_desc: the blit's encode / tone / grade switches and grade strength."""
    _fields_ = [("encode", C.c_uint32), ("tone", C.c_uint32), ("grade", C.c_uint32), ("strength", C.c_float)]


class ReplayReq(C.Structure):
    """vx_replay_req: a camera request of a native replay and its arrival offset."""
    _fields_ = [("t", C.c_double), ("view", CameraView)]


class BatchTuning(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("occupancy", C.c_uint32), ("variant", C.c_uint32),
                ("ordered", C.c_uint32), ("queue_chunk", C.c_uint32), ("waves_per_simd", C.c_uint32),
                ("graph", C.c_uint32)]


class ServeDesc(C.Structure):
    _fields_ = [("lanes", C.c_uint32), ("max_views", C.c_uint32), ("hbm_budget", C.c_uint64),
                ("tuning", BatchTuning)]


class ServeResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int32), ("lane", C.c_uint32), ("batch_views", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("pad", C.c_uint32), ("batch", C.c_uint64),
                ("t_arrival", C.c_double), ("t_submit", C.c_double), ("t_launch", C.c_double),
                ("t_done", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("launches", C.c_uint64),
                ("max_batch_views", C.c_uint64), ("shadow_rays", C.c_uint64), ("pending", C.c_uint64),
                ("used_bytes", C.c_uint64), ("budget_bytes", C.c_uint64)]


class PairDemand(C.Structure):
    _fields_ = [("select", C.c_uint32), ("deferred", C.c_uint32), ("recover", C.c_uint32),
                ("cap", C.c_uint32), ("truncated", C.c_uint32), ("shadow_rays", C.c_uint32)]


class MarchStats(C.Structure):
    _fields_ = [("pixels", C.c_uint64), ("iters", C.c_uint64), ("touched", C.c_uint64), ("capped", C.c_uint64),
                ("worst", C.c_uint32), ("p50", C.c_uint32), ("p90", C.c_uint32), ("p99", C.c_uint32),
                ("top_1", C.c_double), ("top_10", C.c_double), ("chunks", C.c_uint32), ("pad", C.c_uint32),
                ("pairs", PairDemand)]


class Timings(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("tile_select", "march", "hiz", "recover", "march2", "shadow", "total")]


class Hit(C.Structure):
    _fields_ = [("voxel", C.c_int32 * 3), ("face", C.c_uint32), ("material", C.c_uint32),
                ("ci", C.c_uint32), ("t", C.c_float), ("shadow", C.c_uint32)]


HIT_DTYPE = np.dtype([("voxel", "<i4", 3), ("face", "<u4"), ("material", "<u4"), ("ci", "<u4"),
                      ("t", "<f4"), ("shadow", "<u4")])
assert HIT_DTYPE.itemsize == C.sizeof(Hit)

P = C.c_void_p
_sig_host = {
    "vxw_world_new": (P, []),
    "vxw_world_free": (None, [P]),
    "vxw_insert_dense": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, P]),
    "vxw_remove_chunk": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32]),
    "vxw_get_block": (C.c_uint16, [P, C.c_int32, C.c_int32, C.c_int32]),
    "vxw_set_block": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_uint16]),
    "vxw_apply_edits": (C.c_int64, [P, P, C.c_size_t]),
    "vxw_extract_dense": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, P]),
    "vxw_chunk_info_get": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(ChunkInfo)]),
    "vxw_chunk_count": (C.c_size_t, [P]),
    "vxw_chunk_keys": (C.c_size_t, [P, P, C.c_size_t]),
    "vxw_emitter_chunks": (C.c_uint32, [P]),
    "vxw_pool_data": (P, [P, C.c_int]),
    "vxw_pool_bytes": (C.c_size_t, [P, C.c_int]),
    "vxw_take_dirty": (C.c_size_t, [P, C.c_int, P, C.c_size_t]),
    "vxw_pool_stats_get": (C.c_int, [P, C.c_int, C.POINTER(PoolStats)]),
    "vxw_generate": (C.c_int, [P, C.POINTER(GenParams)]),
    "vxw_height": (C.c_int32, [C.POINTER(GenParams), C.c_int32, C.c_int32]),
    "vxw_generate_dense": (C.c_int, [C.POINTER(GenParams), C.c_int32, C.c_int32, C.c_int32, P]),
    "vxw_camera_basis": (None, [C.c_float, C.c_float, P, P, P]),
    "vxw_sun": (None, [C.c_float, P, P]),
    "vxw_build_view": (C.c_int, [P, C.POINTER(Camera), P, P, C.POINTER(View)]),
    "vxw_pack_frame": (C.c_int, [C.POINTER(Camera), C.POINTER(View), C.c_uint32, C.POINTER(Frame)]),
    "vxw_pairs_for": (C.c_uint32, [C.c_uint32, C.c_uint32]),
    "vxw_build_local_flat": (C.c_int, [P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vxw_insert_local_flat": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, P, C.c_size_t]),
    "vxw_insert_dense_key": (C.c_int, [P, C.POINTER(Key), P]),
    "vxw_remove_chunk_key": (C.c_int, [P, C.POINTER(Key)]),
    "vxw_extract_dense_key": (C.c_int, [P, C.POINTER(Key), P]),
    "vxw_chunk_info_key": (C.c_int, [P, C.POINTER(Key), C.POINTER(ChunkInfo)]),
    "vxw_chunk_keys_all": (C.c_size_t, [P, P, C.c_size_t]),
    "vxw_generate_dense_key": (C.c_int, [C.POINTER(GenParams), C.POINTER(Key), P]),
    "vxw_lod_config_default": (None, [C.POINTER(LodConfig)]),
    "vxw_planner_new": (P, []),
    "vxw_planner_free": (None, [P]),
    "vxw_planner_plan": (C.c_int, [P, P, C.POINTER(LodConfig), P, RESIDENCY_FN, P]),
    "vxw_planner_render": (C.c_size_t, [P, P, C.c_size_t]),
    "vxw_planner_desired": (C.c_size_t, [P, P, C.c_size_t]),
    "vxw_build_view_items": (C.c_int, [P, C.POINTER(Camera), P, C.c_size_t, C.POINTER(LodConfig), P, P,
                                       C.POINTER(View)]),
    "vxw_snapshot_new": (P, []),
    "vxw_snapshot_free": (None, [P]),
    "vxw_snapshot_take": (C.c_int, [P, P]),
    "vxw_snapshot_emitter_chunks": (C.c_uint32, [P]),
    "vxw_build_view_snapshot": (C.c_int, [P, C.POINTER(Camera), P, P, C.POINTER(View)]),
    "vxw_snapshot_chunks": (C.c_size_t, [P, P, C.c_size_t]),
    "vxw_snapshot_chunks_all": (C.c_size_t, [P, P, P, C.c_size_t]),
    "vxw_frustum_planes": (C.c_int, [C.POINTER(Camera), P]),
    "vxw_flat_layout": (C.c_int, [P, C.c_size_t, P, C.c_size_t, P, C.c_size_t, P, C.c_uint32, P, P, C.c_size_t,
                                  C.POINTER(C.c_size_t)]),
    "vxw_pack_camera_frame": (C.c_int, [C.POINTER(Camera), C.c_uint32, C.POINTER(Frame)]),
    "vxw_probe_bake": (C.c_int, [C.POINTER(GenParams), P, C.c_int, P, P]),
    "vxw_probe_cube_from_tangents": (None, [P, P]),
    "vxw_probe_bounce_of_uniform_ground": (None, [P, P]),
    "vxw_probe_uploadable": (C.c_int, [P, P]),
    "vxw_surface_block": (C.c_uint16, [C.POINTER(GenParams), C.c_int32, C.c_int32, C.c_int32]),
    "vxw_albedo": (C.c_float, [C.c_uint16, C.c_int]),
    "vxw_shaft_new": (P, [C.c_int32]),
    "vxw_shaft_free": (None, [P]),
    "vxw_shaft_update": (C.c_int, [P, C.POINTER(GenParams), C.c_float, C.c_float, C.c_float]),
    "vxw_shaft_heights": (P, [P]),
    "vxw_shaft_origin_world": (None, [P, P]),
    "vxw_shaft_texel": (C.c_int32, [P]),
    "vxw_shaft_dirty": (C.c_int, [P]),
    "vxw_shaft_envelope": (C.c_float, [P, P, C.c_uint32, C.c_uint32]),
    "vxw_journal_parse": (C.c_int, [P, C.c_size_t, C.c_int32, C.c_int32, C.POINTER(JournalInfo), P, C.c_size_t,
                                    C.c_char_p, C.c_size_t]),
    "vxw_journal_write": (C.c_int, [P, C.c_size_t, C.c_int32, C.c_int32, C.POINTER(Player), P, P, C.c_size_t,
                                    C.POINTER(C.c_size_t)]),
}

_sig_march = {
    "vx_ctx_create": (C.c_int, [C.POINTER(CtxDesc), C.POINTER(P)]),
    "vx_ctx_destroy": (C.c_int, [P]),
    "vx_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
    "vx_get_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
    "vx_resize": (C.c_int, [P, C.c_uint32, C.c_uint32]),
    "vx_pool_sync": (C.c_int, [P, C.c_int, P, C.c_size_t, P, C.c_size_t]),
    "vx_set_chunks": (C.c_int, [P, P, C.c_uint32, P]),
    "vx_pool_mirror_get": (C.c_int, [P, C.c_int, C.POINTER(PoolMirror)]),
    "vx_pool_read": (C.c_int, [P, C.c_int, P, C.c_size_t]),
    "vx_pool_set_flat": (C.c_int, [P, P, C.c_size_t]),
    "vx_taa": (C.c_int, [P]),
    "vx_probe_fill": (C.c_int, [P, P, C.c_int]),
    "vx_probe_write": (C.c_int, [P, P, P, P]),
    "vx_read_probe": (C.c_int, [P, P, C.c_size_t]),
    "vx_shaft_upload": (C.c_int, [P, P, C.c_size_t]),
    "vx_shaft_scan": (C.c_int, [P, C.POINTER(Frame)]),
    "vx_read_shaft": (C.c_int, [P, P, C.c_size_t]),
    "vx_dworld_create": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, P, C.c_uint32, C.POINTER(P)]),
    "vx_dworld_destroy": (C.c_int, [P]),
    "vx_dworld_edit": (C.c_int, [P, P, C.c_size_t]),
    "vx_dworld_compact": (C.c_int, [P]),
    "vx_dworld_chunks": (C.c_int, [P, P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "vx_dworld_extract": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, P]),
    "vx_dworld_stats_get": (C.c_int, [P, C.POINTER(DWorldStats)]),
    "vx_read_taa": (C.c_int, [P, P, C.c_size_t]),
    "vx_write_taa_history": (C.c_int, [P, P, C.c_size_t]),
    "vx_render": (C.c_int, [P, C.POINTER(Frame), C.c_uint32, C.c_uint32]),
    "vx_sync": (C.c_int, [P]),
    "vx_read_vis": (C.c_int, [P, P, C.c_size_t]),
    "vx_read_shadow": (C.c_int, [P, P, C.c_size_t]),
    "vx_read_hiz": (C.c_int, [P, P, C.c_size_t]),
    "vx_write_hiz": (C.c_int, [P, P, C.c_size_t]),
    "vx_read_pairs": (C.c_int, [P, P, C.c_size_t, P]),
    "vx_read_decoded": (C.c_int, [P, P, C.c_size_t]),
    "vx_pair_demand_get": (C.c_int, [P, C.POINTER(PairDemand)]),
    "vx_timings_get": (C.c_int, [P, C.POINTER(Timings)]),
    "vx_timings_reset": (C.c_int, [P]),
    "vx_timings_sum": (C.c_int, [P, C.POINTER(Timings), C.POINTER(C.c_uint32)]),
    "vx_device_outputs": (C.c_int, [P, P, P]),
    "vx_batch_create": (C.c_int, [P, C.c_uint32, C.POINTER(P)]),
    "vx_batch_destroy": (C.c_int, [P]),
    "vx_batch_set_views": (C.c_int, [P, P, C.c_uint32]),
    "vx_batch_set_chunk_table": (C.c_int, [P, P, C.c_uint32]),
    "vx_batch_set_cameras": (C.c_int, [P, P, C.c_uint32]),
    "vx_batch_move_cameras": (C.c_int, [P, P, C.c_uint32]),
    "vx_batch_view_counts": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32)]),
    "vx_batch_read_chunks": (C.c_int, [P, C.c_uint32, P, C.c_size_t, P]),
    "vx_batch_set_tuning": (C.c_int, [P, C.POINTER(BatchTuning)]),
    "vx_batch_get_tuning": (C.c_int, [P, C.POINTER(BatchTuning)]),
    "vx_batch_render": (C.c_int, [P]),
    "vx_batch_sync": (C.c_int, [P]),
    "vx_batch_autotune": (C.c_int, [P, C.c_uint32, C.POINTER(BatchTuning), C.POINTER(C.c_float)]),
    "vx_batch_reset_hiz": (C.c_int, [P]),
    "vx_batch_read_vis": (C.c_int, [P, C.c_uint32, P, C.c_size_t]),
    "vx_batch_read_shadow": (C.c_int, [P, C.c_uint32, P, C.c_size_t]),
    "vx_batch_read_hiz": (C.c_int, [P, C.c_uint32, P, C.c_size_t]),
    "vx_batch_demand_get": (C.c_int, [P, C.POINTER(PairDemand)]),
    "vx_batch_timings_reset": (C.c_int, [P]),
    "vx_batch_timings_sum": (C.c_int, [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint32)]),
    "vx_batch_device_outputs": (C.c_int, [P, C.c_uint32, P, P]),
    "vx_read_heat": (C.c_int, [P, P, C.c_size_t]),
    "vx_resolve_set_assets": (C.c_int, [P, P]),
    "vx_resolve": (C.c_int, [P]),
    "vx_read_color": (C.c_int, [P, P, C.c_size_t]),
    "vx_present": (C.c_int, [P, P]),
    "vx_read_image": (C.c_int, [P, P, C.c_size_t]),
    "vx_march_stats_get": (C.c_int, [P, C.POINTER(MarchStats)]),
    "vx_march_stats_compute": (C.c_int, [P, C.c_size_t, C.c_uint32, C.POINTER(MarchStats)]),
    "vx_batch_set_secondary": (C.c_int, [P, C.c_uint32]),
    "vx_batch_read_secondary": (C.c_int, [P, C.c_uint32, P, C.c_size_t]),
    "vx_batch_secondary_count": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_uint32)]),
    "vx_serve_clock": (C.c_double, []),
    "vx_serve_create": (C.c_int, [P, C.POINTER(ServeDesc), C.POINTER(P)]),
    "vx_serve_destroy": (C.c_int, [P]),
    "vx_serve_submit": (C.c_int, [P, C.c_uint64, C.POINTER(ViewDesc), C.c_double]),
    "vx_serve_set_chunk_table": (C.c_int, [P, P, C.c_uint32]),
    "vx_serve_submit_camera": (C.c_int, [P, C.c_uint64, C.POINTER(CameraView), C.c_double]),
    "vx_serve_poll": (C.c_int, [P, C.POINTER(ServeResult), C.c_uint32, C.POINTER(C.c_uint32), C.c_int]),
    "vx_serve_read_vis": (C.c_int, [P, C.c_uint64, P, C.c_size_t]),
    "vx_serve_read_shadow": (C.c_int, [P, C.c_uint64, P, C.c_size_t]),
    "vx_serve_release": (C.c_int, [P, C.c_uint64]),
    "vx_serve_barrier": (C.c_int, [P]),
    "vx_serve_resume": (C.c_int, [P]),
    "vx_serve_stats_get": (C.c_int, [P, C.POINTER(ServeStats)]),
    "vx_serve_replay_start": (C.c_int, [P, P, C.c_uint32, C.c_uint64, C.c_double, C.c_uint32]),
    "vx_serve_replay_progress": (C.c_int, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "vx_serve_replay_join": (C.c_int, [P, C.POINTER(ServeResult), C.c_uint32]),
    "vx_error_string": (C.c_char_p, [C.c_int]),
    "vx_device_name": (C.c_int, [C.c_char_p, C.c_size_t]),
    "vx_build_local": (C.c_int, [P, P, C.c_uint32, P, C.c_size_t]),
    "vx_build_local_device": (C.c_int, [P, P, C.c_uint32, P, C.c_size_t]),
    "vx_dev_alloc": (C.c_int, [P, C.c_size_t, C.POINTER(P)]),
    "vx_dev_free": (C.c_int, [P, P]),
    "vx_dev_upload": (C.c_int, [P, P, P, C.c_size_t]),
    "vx_dev_download": (C.c_int, [P, P, P, C.c_size_t]),
    "vx_render_noalloc": (C.c_int, [P, C.c_uint32, C.c_uint32, P, P, P, P, P]),
}

_host = None
_march = None


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def check_fresh(lib: str, kind: str):
    """Refuses a library that is missing or was built from other sources than the ones
    beside it (the digest the build wrote next to it differs from the current one)."""
    from . import build as B
    if not os.path.exists(lib):
        raise RuntimeError(f"{lib} is not built; run voxelcraft_menu_amd.build.build()")
    try:
        with open(lib + ".srchash") as fh:
            stamp = fh.read().strip()
    except OSError:
        stamp = ""
    if stamp != B.source_digest(kind):
        raise RuntimeError(f"{lib} does not match its sources (stale or foreign build); "
                           f"rebuild it with voxelcraft_menu_amd.build.build()")


def host():
    """The host world-store library (CPU only)."""
    global _host
    if _host is None:
        check_fresh(HOST_LIB, "host")
        _host = _bind(C.CDLL(HOST_LIB), _sig_host)
    return _host


def march():
    """The HIP ray-march library. Raises if it is not built or stale: there is no CPU fallback."""
    global _march
    if _march is None:
        check_fresh(MARCH_LIB, "march")
        _march = _bind(C.CDLL(MARCH_LIB), _sig_march)
    return _march


_calib = None


def calib():
    """The VALU issue-rate probe (measurement tooling for the benchmark's roofline)."""
    global _calib
    if _calib is None:
        check_fresh(CALIB_LIB, "calib")
        L = C.CDLL(CALIB_LIB)
        L.vxc_valu_probe.restype = C.c_int
        L.vxc_valu_probe.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        _calib = L
    return _calib


def host_symbols():
    return list(_sig_host)


def march_symbols():
    return list(_sig_march)


def check(rc, what="call"):
    if rc < 0:
        raise RuntimeError(f"{what} failed: {ERRORS.get(rc, rc)}")
    return rc


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def march_stats_dict(m: "MarchStats") -> dict:
    out = {k: getattr(m, k) for k, _ in MarchStats._fields_ if k not in ("pairs", "pad")}
    out["pairs"] = {k: int(getattr(m.pairs, k)) for k, _ in PairDemand._fields_}
    return out


def march_stats_of(heat: np.ndarray, chunks: int = 0) -> dict:
    """Renderer::march_stats' summary of a host heat array (vx_march_stats_compute)."""
    a = np.ascontiguousarray(heat, dtype=np.uint32)
    m = MarchStats()
    check(march().vx_march_stats_compute(ptr(a) if a.size else None, a.size, chunks, C.byref(m)),
          "vx_march_stats_compute")
    return march_stats_dict(m)


class DeviceView(C.Structure):
    """vx_device_view: device-resident inputs of vx_render_noalloc."""
    _fields_ = [("inners", C.c_void_p), ("inners_bytes", C.c_uint64),
                ("leaves", C.c_void_p), ("leaves_bytes", C.c_uint64),
                ("bricks", C.c_void_p), ("bricks_bytes", C.c_uint64),
                ("chunks", C.c_void_p), ("n_chunks", C.c_uint32),
                ("grid", C.c_void_p)]


NOALLOC_MAX_TILES = 1 << 17

# vx_sec_hit: one secondary leg (t, voxel, info = ci | normal << 13 | traced << 30 | hit << 31, block id)
SEC_DTYPE = np.dtype([("t", "<f4"), ("voxel", "<u4"), ("info", "<u4"), ("material", "<u4")])
SEC_SUN_BLOCKED, SEC_SUN_TESTED, SEC_TRACED, SEC_HIT = 1 << 28, 1 << 29, 1 << 30, 1 << 31
