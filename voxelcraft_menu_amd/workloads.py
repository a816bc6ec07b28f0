"""The benchmark workloads: worlds, cameras and flag words.

Five workloads, indexed 0..4:
  0  plain march, small: 256x256x128 world (chunks x,z in [0,4), y in {1,2}), seed 1,
     one 640x360 camera at a fixed pose, FOV 70, primary + 1 sun shadow ray per pixel
  1  plain march at scale: 1024x1024x256 world (x,z in [0,16), y in [0,4)), seed 7,
     caves below the surface, ~2% leaves/glass on the surface, one 1920x1080 camera,
     primary + sun shadow, two-phase Hi-Z on 8x8 tiles
  2  batching: the workload-1 world, 64 views of uneven resolution drawn uniformly from
     {256x144, 640x360, 1280x720, 1920x1080}, random poses above the terrain (seed 2)
  3  incoherent batch: 2048x2048x512 world (x,z in [0,32), y in [0,8)), sea level 96,
     ~40% occupancy underground, glass 1%; 32 views at 1280x720 (seed 3)
  4  serving: the workload-3 world resident on the device, Poisson view requests with
     workload-2 resolutions (seed 4) and 1000-voxel edit batches applied in place

Default camera (src/config.rs:274-393): FOV 70 deg, yaw 40, pitch -14, 40 blocks above
the ground plus eye height 1.62 (src/player.rs:10), time of day 0.30 for the sun
(src/scene.rs:10-15), view distance 2600 (src/lod.rs:44). Foliage tufts are off: they
sample the texture atlas with hardware filtering and are not part of this path.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _abi as A
from .world import CameraParams, GenParams, World

EYE = 1.62

# Reference default flag word (src/scene.rs:422-542 with the src/config.rs defaults)
# minus FLAG_FOLIAGE.
DEFAULT_FLAGS = (A.FLAG_SHADOWS | A.FLAG_AO | A.FLAG_HIZ | A.FLAG_WATER_REFLECT | A.FLAG_WATER_REFRACT
                 | A.FLAG_TINT | 512 | 1024 | 2048 | 4096 | 8192 | 16384 | A.FLAG_LEAF_CUTOUT | 1048576
                 | A.FLAG_WATER_FAR | A.FLAG_WATER_DARK | A.FLAG_SNELL | 67108864 | 134217728 | 536870912
                 | 268435456 | 2147483648 | A.FLAG_GLASS | A.FLAG_DISTANT_SHADOWS | A.FLAG_WATER_SHADOW_CUT
                 | A.FLAG_LEAF_THIN)
DEFAULT_SPEC_HI = A.FLAG_HI_SKY_SPECULAR | A.FLAG_HI_FULL_MARCH | A.FLAG_HI_SHADOW_PASS | 8 | 16 | 32

# Pipeline-key masks (src/render/mod.rs:272-353).
SPEC_MASK = (128 | 256 | 512 | 1024 | 2048 | 4096 | 8192 | 16384 | 32768 | 65536 | 131072 | 262144 | 524288
             | 1048576 | 16777216 | 33554432 | 67108864 | 134217728 | 268435456 | 536870912 | 1073741824
             | 2147483648)

RESOLUTIONS = [(256, 144), (640, 360), (1280, 720), (1920, 1080)]


@dataclass
class Workload:
    index: int
    name: str
    gen: GenParams
    width: int
    height: int
    views: int = 1


WORKLOADS = {
    0: Workload(0, "plain-small-256x256x128-640x360",
                GenParams(seed=1, sea_level=128, chunk_min=(0, 1, 0), chunk_max=(4, 3, 4), tree_density=0.004),
                640, 360),
    1: Workload(1, "plain-1024x1024x256-1920x1080",
                GenParams(seed=7, sea_level=128, chunk_min=(0, 0, 0), chunk_max=(16, 4, 16), cave_fraction=0.65,
                          tree_density=0.004, glass_fraction=0.01, leaf_extra=0.01),
                1920, 1080),
    2: Workload(2, "batch-64-views-1024x1024x256",
                GenParams(seed=7, sea_level=128, chunk_min=(0, 0, 0), chunk_max=(16, 4, 16), cave_fraction=0.65,
                          tree_density=0.004, glass_fraction=0.01, leaf_extra=0.01),
                0, 0, views=64),
    3: Workload(3, "incoherent-32x1280x720-2048x2048x512",
                GenParams(seed=7, sea_level=96, chunk_min=(0, 0, 0), chunk_max=(32, 8, 32), cave_fraction=0.60,
                          tree_density=0.004, glass_fraction=0.01, leaf_extra=0.0),
                1280, 720, views=32),
    4: Workload(4, "serving-2048x2048x512",
                GenParams(seed=7, sea_level=96, chunk_min=(0, 0, 0), chunk_max=(32, 8, 32), cave_fraction=0.60,
                          tree_density=0.004, glass_fraction=0.01, leaf_extra=0.0),
                0, 0),
}


def make_world(w: Workload, threads: int = 0) -> World:
    world = World()
    g = w.gen
    g.threads = threads
    world.generate(g)
    return world


def ground(w: Workload, x: int, z: int) -> int:
    return max(w.gen.height(x, z), w.gen.sea_level)


def default_camera(w: Workload, flags: int = DEFAULT_FLAGS, width: int = 0, height: int = 0) -> CameraParams:
    """World-centre column, 40 above the ground, yaw 40, pitch -14 (src/config.rs:383-387)."""
    sx = (w.gen.chunk_max[0] - w.gen.chunk_min[0]) * 64
    sz = (w.gen.chunk_max[2] - w.gen.chunk_min[2]) * 64
    cx = w.gen.chunk_min[0] * 64 + sx // 2
    cz = w.gen.chunk_min[2] * 64 + sz // 2
    y = ground(w, cx, cz) + 40.0 + EYE
    return CameraParams(pos=(cx + 0.5, y, cz + 0.5), yaw=math.radians(40.0), pitch=math.radians(-14.0),
                        width=width or w.width, height=height or w.height, sea_level=float(w.gen.sea_level),
                        flags=flags)


def random_cameras(w: Workload, n: int, seed: int, resolutions=None, flags: int = DEFAULT_FLAGS):
    """Seeded poses above the terrain: random column, 20-60 above ground, yaw U[0,360),
    pitch U[-30, 5] degrees; resolution uniform over the list."""
    rng = np.random.default_rng(seed)
    res = resolutions or RESOLUTIONS
    sx = (w.gen.chunk_max[0] - w.gen.chunk_min[0]) * 64
    sz = (w.gen.chunk_max[2] - w.gen.chunk_min[2]) * 64
    cams = []
    for _ in range(n):
        x = int(w.gen.chunk_min[0] * 64 + rng.integers(8, sx - 8))
        z = int(w.gen.chunk_min[2] * 64 + rng.integers(8, sz - 8))
        up = float(rng.uniform(20.0, 60.0))
        yaw = float(rng.uniform(0.0, 360.0))
        pitch = float(rng.uniform(-30.0, 5.0))
        rw, rh = res[int(rng.integers(0, len(res)))]
        y = ground(w, x, z) + up + EYE
        cams.append(CameraParams(pos=(x + 0.5, y, z + 0.5), yaw=math.radians(yaw), pitch=math.radians(pitch),
                                 width=rw, height=rh, sea_level=float(w.gen.sea_level), flags=flags))
    return cams


def spec_words(flags: int, world: World, spec_hi: int = DEFAULT_SPEC_HI):
    """(spec, spec_hi) pipeline key as Renderer::render forms it (src/render/mod.rs:1597,
    emitter bit from src/headless.rs:464-468)."""
    hi = spec_hi | (A.FLAG_HI_EMITTER_WORLD if world.any_emitter_resident() else 0)
    return flags & SPEC_MASK, hi
