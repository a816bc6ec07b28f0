"""ctypes binding of the CPU oracle (oracle/vxoracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product never imports it.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libvxoracle.so")


class Pools(C.Structure):
    _fields_ = [("inners", C.c_void_p), ("n_inners", C.c_size_t),
                ("leaves", C.c_void_p), ("n_leaves", C.c_size_t),
                ("bricks", C.c_void_p), ("n_bricks", C.c_size_t)]


class Spec(C.Structure):
    _fields_ = [("leaf_cutout", C.c_uint32), ("glass", C.c_uint32), ("full_march", C.c_uint32),
                ("emitter_gather", C.c_uint32), ("water_repair", C.c_uint32), ("leaf_fill", C.c_float),
                ("glass_reflect", C.c_uint32), ("waves", C.c_uint32), ("wave_aniso", C.c_uint32),
                ("wave_shoal", C.c_uint32), ("wave_fill", C.c_uint32), ("snell_bend", C.c_uint32)]


_STAT_FIELDS = ["pairs_select", "pairs_deferred", "pairs_recover", "rays_primary", "rays_shadow",
                "chunk_marches", "iters", "l1_loads", "leaf_loads", "attr_loads", "brick_words",
                "grid_loads", "chunk_loads", "node_bytes", "capped", "oob", "rays_secondary", "rays_secondary_sun"]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in _STAT_FIELDS]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n in _STAT_FIELDS}


class Hit(C.Structure):
    _fields_ = [("hit", C.c_uint32), ("t", C.c_float), ("normal", C.c_uint32),
                ("voxel", C.c_uint32 * 3), ("micro", C.c_uint32 * 3), ("iters", C.c_uint32)]


SEC = np.dtype([("t", "<f4"), ("voxel", "<u4"), ("info", "<u4"), ("material", "<u4")])
DECODED = np.dtype([("voxel", "<i4", 3), ("face", "<u4"), ("material", "<u4"), ("ci", "<u4"), ("t", "<f4")])

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with the recipe in oracle/Makefile."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB


def _bind(L):
    V = C.c_void_p
    L.vxo_render.restype = C.c_int
    L.vxo_render.argtypes = [C.POINTER(Pools), V, C.c_uint32, V, V, C.POINTER(Spec), V, V, V,
                             C.POINTER(Stats), C.c_int]
    L.vxo_render_heat.restype = C.c_int
    L.vxo_render_heat.argtypes = [C.POINTER(Pools), V, C.c_uint32, V, V, C.POINTER(Spec), V, V, V, V,
                                  C.POINTER(Stats), C.c_int]
    L.vxo_secondary.restype = C.c_int
    L.vxo_secondary.argtypes = [C.POINTER(Pools), V, C.c_uint32, V, V, C.POINTER(Spec), V, V,
                                C.POINTER(Stats), C.c_int]
    L.vxo_primary_shadow.restype = C.c_int
    L.vxo_primary_shadow.argtypes = [C.POINTER(Pools), V, C.c_uint32, V, V, C.POINTER(Spec), V, V,
                                     C.POINTER(Stats), C.c_int]
    L.vxo_decode.restype = C.c_int
    L.vxo_decode.argtypes = [C.POINTER(Pools), V, C.c_uint32, V, V, V, C.c_int]
    L.vxo_ray_dir.restype = None
    L.vxo_ray_dir.argtypes = [V, C.c_float, C.c_float, V]
    L.vxo_depth_key.restype = C.c_uint32
    L.vxo_depth_key.argtypes = [V, C.c_float]
    L.vxo_make_key.restype = C.c_uint64
    L.vxo_make_key.argtypes = [V, C.c_float, C.c_uint32, C.c_uint32, V, V]
    L.vxo_hash_u32.restype = C.c_uint32
    L.vxo_hash_u32.argtypes = [C.c_uint32]
    L.vxo_f32_to_f16.restype = C.c_uint16
    L.vxo_f32_to_f16.argtypes = [C.c_float]
    L.vxo_block_at.restype = C.c_uint32
    L.vxo_block_at.argtypes = [C.POINTER(Pools), V, C.c_uint32, V]
    L.vxo_march_chunk.restype = None
    L.vxo_march_chunk.argtypes = [C.POINTER(Pools), V, C.c_uint32, V, V, C.POINTER(Spec), C.c_uint32,
                                  V, V, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Hit)]
    L.vxo_shadow_ray_t.restype = C.c_float
    L.vxo_shadow_ray_t.argtypes = [C.POINTER(Pools), V, C.c_uint32, V, V, C.POINTER(Spec), V, V,
                                   C.c_float]
    L.vxo_hardware_threads.restype = C.c_int
    L.vxo_wave_cos.restype = C.c_float
    L.vxo_wave_cos.argtypes = [C.c_float]
    L.vxo_wave_exp.restype = C.c_float
    L.vxo_wave_exp.argtypes = [C.c_float]
    L.vxo_water_reflect_dir.restype = None
    L.vxo_water_reflect_dir.argtypes = [C.POINTER(Pools), V, C.c_uint32, V, V, C.POINTER(Spec), C.c_uint32, V, V,
                                        C.c_float, V]
    L.vxo_log2.restype = C.c_float
    L.vxo_log2.argtypes = [C.c_float]
    L.vxo_exp2.restype = C.c_float
    L.vxo_exp2.argtypes = [C.c_float]
    L.vxo_pow.restype = C.c_float
    L.vxo_pow.argtypes = [C.c_float, C.c_float]
    L.vxo_resolve.restype = C.c_int
    L.vxo_resolve.argtypes = [C.POINTER(Pools), V, C.c_uint32, V, V, C.POINTER(RSpec), C.POINTER(ResolveAssets),
                              V, V, V, V, C.POINTER(Stats), C.c_int]
    L.vxo_shaft_scan.restype = C.c_int
    L.vxo_shaft_scan.argtypes = [V, V, V]
    L.vxo_terrain_shade.restype = C.c_float
    L.vxo_terrain_shade.argtypes = [V, V, V, C.c_float]
    L.vxo_soft_mask.restype = C.c_int
    L.vxo_soft_mask.argtypes = [C.POINTER(Pools), V, C.c_uint32, V, V, C.POINTER(RSpec), V, V, V, V, C.c_int]
    L.vxo_probe_tap.restype = None
    L.vxo_probe_tap.argtypes = [V, V, V]
    L.vxo_probe_field.restype = None
    L.vxo_probe_field.argtypes = [V, V, C.c_uint32, V]
    L.vxo_biome_noise.restype = C.c_float
    L.vxo_biome_noise.argtypes = [C.c_int32, C.c_float, C.c_float, C.c_float]
    L.vxo_biome_tint.restype = None
    L.vxo_biome_tint.argtypes = [V, C.c_uint32, C.c_float, C.c_float, V]
    L.vxo_taa.restype = C.c_int
    L.vxo_taa.argtypes = [C.POINTER(Pools), V, C.c_uint32, V, V, V, V, V]
    L.vxo_present.restype = None
    L.vxo_present.argtypes = [V, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, V]
    L.vxo_wave_field.restype = None
    L.vxo_wave_field.argtypes = [V, C.POINTER(Spec), C.c_float, C.c_float, C.c_float, V, C.c_float, C.c_int,
                                 C.c_float, V]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _bind(C.CDLL(LIB))
    return _lib


@contextlib.contextmanager
def planted(n: int, tmpdir: str):
    """Every oracle call inside the block runs a copy compiled with planted mistake n
    (VXO_PLANT in vxoracle.c), built into tmpdir: for tests showing that a check rejects
    a wrong restatement."""
    global _lib
    out = os.path.join(str(tmpdir), f"libvxoracle_plant{n}.so")
    subprocess.run(["make", "-C", HERE, "-s", "-B", f"PLANT={n}", f"OUT={out}"], check=True)
    saved = lib()
    _lib = _bind(C.CDLL(out))
    try:
        yield _lib
    finally:
        _lib = saved


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class RSpec(C.Structure):
    """vxo_rspec: the resolve's override switches; `lanes` adds the glass and water lanes,
    whose traces read the march switches `march` points to."""
    _fields_ = [(k, C.c_uint32) for k in ("sky_specular", "lighting_repair", "sky_tint", "sky_cool", "sky_look",
                                          "light_rgb", "emitter_gather", "shore_wet", "tex_variation", "glass",
                                          "lanes", "water_sec", "shore_foam", "flat_secondary")] + \
        [("march", C.c_void_p), ("distant", C.c_uint32), ("shaft", C.c_void_p), ("tint", C.c_uint32),
         ("tint_balance", C.c_uint32), ("probe_cube", C.c_uint32), ("probe_bounce", C.c_uint32),
         ("probe_sun", C.c_uint32), ("probe_sun_high", C.c_uint32), ("probe", C.c_void_p),
         ("caustics", C.c_uint32), ("soft_shadows", C.c_uint32), ("probe_tap", C.c_uint32),
         ("probe_ambient", C.c_uint32)]

    def with_march(self, spec: "Spec") -> "RSpec":
        self._march = spec  # kept alive with the record
        self.march = C.cast(C.pointer(spec), C.c_void_p).value
        return self

    def with_probe(self, field: np.ndarray) -> "RSpec":
        """The probe field (PROBE_SHAPE f16 bit patterns) the cube and bounce terms sample."""
        self._probe = np.ascontiguousarray(field, dtype=np.uint16)
        assert self._probe.shape == PROBE_SHAPE
        self.probe = self._probe.ctypes.data
        return self

    def with_shaft(self, result: np.ndarray) -> "RSpec":
        """Distant terrain shadows from a shaft_scan result slice (512 x 512 f32)."""
        self._shaft = np.ascontiguousarray(result, dtype=np.float32).reshape(-1)
        assert self._shaft.size == SHAFT_DIM * SHAFT_DIM
        self.shaft = self._shaft.ctypes.data
        self.distant = 1
        return self


class ResolveAssets(C.Structure):
    _fields_ = [("atlas", C.c_void_p), ("size", C.c_uint32), ("layers", C.c_uint32), ("mips", C.c_uint32),
                ("block_faces", C.c_void_p), ("n_faces", C.c_uint32)]


SHAFT_DIM = 512  # src/shaft.rs:3
PROBE_SHAPE = (128, 768, 128, 4)  # probe_tex (z, y, x, rgba), src/render/mod.rs:2240-2268
PROBE_BITS = {"cube": 67108864, "bounce": 134217728, "sun": 536870912, "sun_high": 1073741824,
              "tap": 16777216, "ambient": 33554432}
SHAFT_SOFT = 4.0  # src/shaft.rs:9


def rspec_from(flags: int, spec_hi: int, shaft: np.ndarray | None = None, probe: np.ndarray | None = None) -> RSpec:
    """The resolve's override values for a (flags, spec_hi) pipeline key (src/render/mod.rs:
    2355-2689). A key with FLAG_DISTANT_SHADOWS needs the shaft field's scan result, one
    with FLAG_PROBE_CUBE or FLAG_PROBE_BOUNCE the probe field."""
    distant = bool(flags & 32768)
    if distant and shaft is None:
        raise ValueError("distant terrain shadows read the shaft field: pass its scan result")
    sampled = bool(flags & (PROBE_BITS["cube"] | PROBE_BITS["bounce"] | PROBE_BITS["tap"] | PROBE_BITS["ambient"]))
    if sampled and probe is None:
        raise ValueError("the probe cube, bounce, tap and ambient read the probe field: pass it")
    r = _rspec(flags, spec_hi)
    for k, bit in PROBE_BITS.items():
        setattr(r, f"probe_{k}", 1 if flags & bit else 0)
    if sampled:
        r.with_probe(probe)
    return r.with_shaft(shaft) if distant else r


def probe_field(field: np.ndarray, p, nid: int) -> np.ndarray:
    """probe_field (resolve.wgsl:106-118) of a PROBE_SHAPE field at world p: rgba."""
    fl = np.ascontiguousarray(field, dtype=np.uint16)
    q = np.asarray(p, dtype=np.float32)
    out = np.zeros(4, np.float32)
    lib().vxo_probe_field(_p(fl), _p(q), nid, _p(out))
    return out


def probe_tap(field: np.ndarray, p) -> np.ndarray:
    """probe_tap (resolve.wgsl:120-127) of a PROBE_SHAPE field at world p: rgb."""
    fl = np.ascontiguousarray(field, dtype=np.uint16)
    q = np.asarray(p, dtype=np.float32)
    out = np.zeros(3, np.float32)
    lib().vxo_probe_tap(_p(fl), _p(q), _p(out))
    return out


def shaft_scan(heights: np.ndarray, frame_bytes: bytes) -> np.ndarray:
    """shaft_scan's eight passes (shaft.wgsl) over a 512 x 512 height field with the frame's
    sun_dir and shaft_texel: the SHAFT_RESULT slice, (512, 512) f32."""
    hgt = np.ascontiguousarray(heights, dtype=np.float32).reshape(-1)
    assert hgt.size == SHAFT_DIM * SHAFT_DIM
    f = np.frombuffer(frame_bytes, dtype=np.uint8).copy()
    out = np.zeros((SHAFT_DIM, SHAFT_DIM), dtype=np.float32)
    if lib().vxo_shaft_scan(_p(hgt), _p(f), _p(out)) != 0:
        raise RuntimeError("oracle shaft scan rejected its arguments")
    return out


def terrain_shade(result: np.ndarray, frame_bytes: bytes, p, d: float = 0.0) -> float:
    """terrain_shade(p + sun_dir * d) (common.wgsl:1571-1580) over a scan result."""
    r = np.ascontiguousarray(result, dtype=np.float32).reshape(-1)
    f = np.frombuffer(frame_bytes, dtype=np.uint8).copy()
    q = np.asarray(p, dtype=np.float32)
    return float(lib().vxo_terrain_shade(_p(r), _p(f), _p(q), d))


def biome_noise(seed: int, freq: float, x: float, z: float) -> float:
    """biome_noise (resolve.wgsl:201-257) at world (x, z)."""
    return float(lib().vxo_biome_noise(seed, freq, x, z))


def biome_tint(frame_bytes: bytes, x: float, z: float, balance: int = 0) -> np.ndarray:
    """biome_tint (resolve.wgsl:274-287) at world (x, z) with the frame's tint fields."""
    f = np.frombuffer(frame_bytes, dtype=np.uint8).copy()
    out = np.zeros(3, np.float32)
    lib().vxo_biome_tint(_p(f), balance, x, z, _p(out))
    return out


def _rspec(flags: int, spec_hi: int) -> RSpec:
    return RSpec(sky_specular=1 if spec_hi & 1 else 0, lighting_repair=0 if spec_hi & 1048576 else 1,
                 sky_tint=1 if flags & 2147483648 else 0, sky_cool=1 if spec_hi & 2048 else 0,
                 sky_look=1 if spec_hi & 1024 else 0, light_rgb=1 if flags & 268435456 else 0,
                 emitter_gather=1 if spec_hi & 4 else 0, shore_wet=1 if spec_hi & 8 else 0,
                 tex_variation=1 if flags & 16384 else 0, glass=1 if flags & 131072 else 0,
                 lanes=1, water_sec=1 if spec_hi & 32 else 0, shore_foam=1 if spec_hi & 16 else 0,
                 flat_secondary=1 if flags & 1048576 else 0, tint=1 if flags & 128 else 0,
                 tint_balance=1 if spec_hi & 512 else 0, caustics=1 if spec_hi & 64 else 0,
                 soft_shadows=1 if spec_hi & 65536 else 0).with_march(
                     spec_from(flags, spec_hi))


def resolve(scene: "Scene", frame_bytes: bytes, rspec: RSpec, atlas: np.ndarray, size: int, layers: int, mips: int,
            block_faces: np.ndarray, vis: np.ndarray, sun: np.ndarray | None, heat: np.ndarray | None = None,
            threads: int = 0) -> np.ndarray:
    """The resolve of a frame: (H, W, 4) float32; lanes 1 and 2 (glass and water pixels) when
    rspec.lanes is set, else alpha 0 there."""
    f = np.frombuffer(frame_bytes, dtype=np.uint8).copy()
    res = np.frombuffer(frame_bytes[80:88], dtype=np.uint32)
    w, h = int(res[0]), int(res[1])
    at = np.ascontiguousarray(atlas, dtype=np.uint8).reshape(-1)
    fc = np.ascontiguousarray(block_faces, dtype=np.uint32).reshape(-1)
    a = ResolveAssets(at.ctypes.data, size, layers, mips, fc.ctypes.data, fc.size)
    v = np.ascontiguousarray(vis, dtype=np.uint64)
    s = None if sun is None else np.ascontiguousarray(sun, dtype=np.uint16)
    hz = None if heat is None else np.ascontiguousarray(heat, dtype=np.uint32)
    out = np.zeros((h, w, 4), dtype=np.float32)
    st = Stats()
    rc = lib().vxo_resolve(C.byref(scene.pools), _p(scene.chunks), scene.n_chunks, _p(scene.grid), _p(f),
                           C.byref(rspec), C.byref(a), _p(v), _p(s), _p(hz), _p(out), C.byref(st), threads)
    if rc != 0:
        raise RuntimeError("oracle resolve rejected its arguments")
    return out


def soft_mask(scene: "Scene", frame_bytes: bytes, rspec: RSpec, vis: np.ndarray, sun: np.ndarray,
              threads: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """shadow_horizontal and shadow_vertical (shadow.wgsl:49-116) alone: the row pass's
    (H, W, 2) f16 (occlusion, support) and the (H, W) u8 mask of clamp(1 - occlusion)."""
    f = np.frombuffer(frame_bytes, dtype=np.uint8).copy()
    res = np.frombuffer(frame_bytes[80:88], dtype=np.uint32)
    w, h = int(res[0]), int(res[1])
    v = np.ascontiguousarray(vis, dtype=np.uint64)
    s = np.ascontiguousarray(sun, dtype=np.uint16)
    hpass = np.zeros((h, w, 2), dtype=np.uint16)
    mask = np.zeros((h, w), dtype=np.uint8)
    rc = lib().vxo_soft_mask(C.byref(scene.pools), _p(scene.chunks), scene.n_chunks, _p(scene.grid), _p(f),
                             C.byref(rspec), _p(v), _p(s), _p(hpass), _p(mask), threads)
    if rc != 0:
        raise RuntimeError("oracle soft mask rejected its arguments")
    return hpass.view(np.float16), mask


def taa(scene: "Scene", frame_bytes: bytes, vis: np.ndarray, src: np.ndarray, hist: np.ndarray) -> np.ndarray:
    """The TAA pass (taa.wgsl): src the resolve colour (H, W, 4) f32, hist the previous
    output (H, W, 4) f16 bits. Returns this frame's output, (H, W, 4) f16 bits."""
    f = np.frombuffer(frame_bytes, dtype=np.uint8).copy()
    res = np.frombuffer(frame_bytes[80:88], dtype=np.uint32)
    w, h = int(res[0]), int(res[1])
    s = np.ascontiguousarray(src, dtype=np.float32).reshape(h, w, 4)
    hs = np.ascontiguousarray(hist, dtype=np.uint16).reshape(h, w, 4)
    v = np.ascontiguousarray(vis, dtype=np.uint64).reshape(-1)
    out = np.zeros((h, w, 4), dtype=np.uint16)
    rc = lib().vxo_taa(C.byref(scene.pools), _p(scene.chunks), scene.n_chunks, _p(f), _p(v), _p(s), _p(hs), _p(out))
    if rc != 0:
        raise RuntimeError(f"vxo_taa failed ({rc})")
    return out


def present(rgba: np.ndarray, encode: int = 1, tone: int = 0, grade: int = 0, strength: float = 1.0) -> np.ndarray:
    """The present blit of an (H, W, 4) float32 colour: (H, W, 4) uint8."""
    src = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = src.shape[:2]
    out = np.zeros((h, w, 4), dtype=np.uint8)
    lib().vxo_present(_p(src), w, h, encode, tone, grade, strength, _p(out))
    return out


def spec_from(flags: int, spec_hi: int) -> Spec:
    """WGSL override values for a (flags, spec_hi) pipeline key (src/render/mod.rs:2355-2689)."""
    return Spec(
        leaf_cutout=1 if flags & 65536 else 0,
        glass=1 if flags & 131072 else 0,
        full_march=1 if spec_hi & 2 else 0,
        emitter_gather=1 if spec_hi & 4 else 0,
        water_repair=0 if spec_hi & 2097152 else 1,
        leaf_fill=0.62 if flags & 524288 else 0.72,
        glass_reflect=1 if spec_hi & 128 else 0,
        waves=1 if flags & 512 else 0,
        wave_aniso=1 if flags & 1024 else 0,
        wave_shoal=1 if flags & 2048 else 0,
        wave_fill=1 if flags & 4096 else 0,
        snell_bend=1 if spec_hi & 256 else 0,
    )


class Scene:
    """Pools + chunk list + grid as numpy arrays, kept alive for the oracle calls."""

    def __init__(self, inners: np.ndarray, leaves: np.ndarray, bricks: np.ndarray,
                 chunks: np.ndarray, grid: np.ndarray):
        self.inners = np.ascontiguousarray(inners, dtype=np.uint32).reshape(-1)
        self.leaves = np.ascontiguousarray(leaves, dtype=np.uint32).reshape(-1)
        self.bricks = np.ascontiguousarray(bricks, dtype=np.uint32).reshape(-1)
        self.chunks = np.ascontiguousarray(chunks)
        self.n_chunks = self.chunks.nbytes // 96
        self.grid = np.ascontiguousarray(grid, dtype=np.uint32)
        self.pools = Pools(_p(self.inners) if self.inners.size else None, self.inners.size // 4,
                           _p(self.leaves) if self.leaves.size else None, self.leaves.size // 2,
                           _p(self.bricks) if self.bricks.size else None, self.bricks.size)


def render(scene: Scene, frame_bytes: bytes, spec: Spec, hiz: np.ndarray, shadow: bool = True,
           threads: int = 0):
    """Run the reference pass sequence on the CPU. Returns (vis, shadow_f16, hiz_out, stats)."""
    f = np.frombuffer(frame_bytes, dtype=np.uint8).copy()
    res = np.frombuffer(frame_bytes[80:88], dtype=np.uint32)
    w, h = int(res[0]), int(res[1])
    vis = np.zeros(w * h, dtype=np.uint64)
    sh = np.zeros(w * h, dtype=np.uint16) if shadow else None
    hz = np.ascontiguousarray(hiz, dtype=np.uint32).copy()
    st = Stats()
    rc = lib().vxo_render(C.byref(scene.pools), _p(scene.chunks), scene.n_chunks, _p(scene.grid), _p(f),
                          C.byref(spec), _p(hz), _p(vis), _p(sh), C.byref(st), threads)
    if rc != 0:
        raise RuntimeError("oracle render rejected its arguments")
    return vis, sh, hz, st.as_dict()


def render_heat(scene: Scene, frame_bytes: bytes, spec: Spec, hiz: np.ndarray, shadow: bool = True,
                threads: int = 0):
    """render() plus the FLAG_HEATMAP counter (march.wgsl:51-53): per pixel, the sum of
    march_chunk iterations over every marched pair. Returns (vis, shadow, hiz_out, heat, stats)."""
    f = np.frombuffer(frame_bytes, dtype=np.uint8).copy()
    res = np.frombuffer(frame_bytes[80:88], dtype=np.uint32)
    w, h = int(res[0]), int(res[1])
    vis = np.zeros(w * h, dtype=np.uint64)
    heat = np.zeros(w * h, dtype=np.uint32)
    sh = np.zeros(w * h, dtype=np.uint16) if shadow else None
    hz = np.ascontiguousarray(hiz, dtype=np.uint32).copy()
    st = Stats()
    rc = lib().vxo_render_heat(C.byref(scene.pools), _p(scene.chunks), scene.n_chunks, _p(scene.grid), _p(f),
                               C.byref(spec), _p(hz), _p(vis), _p(sh), _p(heat), C.byref(st), threads)
    if rc != 0:
        raise RuntimeError("oracle render rejected its arguments")
    return vis, sh, hz, heat, st.as_dict()


def secondary(scene: Scene, frame_bytes: bytes, spec: Spec, vis: np.ndarray, threads: int = 0):
    """Secondary legs of every pixel: (W*H, 4) array of SEC records ([:, 0] refract or
    transmit, [:, 1] reflect, [:, 2:4] the refract and reflect legs from where a glass
    transmit leg lands on water), and the stats of the pass."""
    f = np.frombuffer(frame_bytes, dtype=np.uint8).copy()
    v = np.ascontiguousarray(vis, dtype=np.uint64)
    out = np.zeros((v.size, 4), dtype=SEC)
    st = Stats()
    rc = lib().vxo_secondary(C.byref(scene.pools), _p(scene.chunks), scene.n_chunks, _p(scene.grid), _p(f),
                             C.byref(spec), _p(v), _p(out), C.byref(st), threads)
    if rc != 0:
        raise RuntimeError("oracle secondary rejected its arguments")
    return out, st.as_dict()


def primary_shadow(scene: Scene, frame_bytes: bytes, spec: Spec, vis: np.ndarray, threads: int = 0):
    f = np.frombuffer(frame_bytes, dtype=np.uint8).copy()
    sh = np.zeros(vis.size, dtype=np.uint16)
    st = Stats()
    v = np.ascontiguousarray(vis, dtype=np.uint64)
    rc = lib().vxo_primary_shadow(C.byref(scene.pools), _p(scene.chunks), scene.n_chunks, _p(scene.grid), _p(f),
                                  C.byref(spec), _p(v), _p(sh), C.byref(st), threads)
    if rc != 0:
        raise RuntimeError("oracle shadow rejected its arguments")
    return sh, st.as_dict()


def decode(scene: Scene, frame_bytes: bytes, vis: np.ndarray):
    f = np.frombuffer(frame_bytes, dtype=np.uint8).copy()
    out = np.zeros(vis.size, dtype=DECODED)
    v = np.ascontiguousarray(vis, dtype=np.uint64)
    lib().vxo_decode(C.byref(scene.pools), _p(scene.chunks), scene.n_chunks, _p(f), _p(v), _p(out), 1)
    return out


def hardware_threads() -> int:
    return int(lib().vxo_hardware_threads())
