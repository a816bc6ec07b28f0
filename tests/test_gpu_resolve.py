"""GPU parity of the resolve, lane 0 (vx_resolve), against the oracle through the C ABI.

Every pixel's RGBA32F colour must equal the oracle's bit for bit (both run the same f32
steps: the library's own log2 / exp2 / pow / exp, the atlas filtering restated): on the
floor-and-pillar chunk of the closed-form pins under each lighting switch, on views of
the 256x256x128 world (water, leaves, glass and sand), with fog and the sky look, for a
heatmap frame, and after a second frame with Hi-Z carried. Views that need what is not
restated are refused.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from tests.test_oracle_resolve_pins import LAYERS, MIPS, SIZE, _faces
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.renderer import Renderer
from voxelcraft_menu_amd.world import World, build_view

pytestmark = pytest.mark.gpu

SKY_TINT_BIT, LIGHT_RGB_BIT, TEX_VARIATION_BIT, TINT_BIT = 2147483648, 268435456, 16384, 128
SKY_SPECULAR, EMITTER, SHORE_WET, SKY_LOOK, SKY_COOL, LEGACY_LIGHTING = 1, 4, 8, 1024, 2048, 1048576
PROBES = 16777216 | 33554432 | 67108864 | 134217728 | 536870912 | 1073741824
# the workload flags without probes (not restated) and without the biome tint and distant
# shadows, whose own tests set up their frame fields
BASE_FLAGS = W.DEFAULT_FLAGS & ~(PROBES | TINT_BIT | 32768)


def _atlas(seed=1):
    rng = np.random.default_rng(seed)
    n = sum(LAYERS * (SIZE >> l) ** 2 for l in range(MIPS))
    return rng.integers(0, 256, n * 4, dtype=np.uint8)


def _faces_perm():
    f = _faces()
    return f | ((np.arange(f.size, dtype=np.uint32).reshape(f.shape) % 3) << 16)


@pytest.fixture(scope="module")
def small_world():
    return W.make_world(W.WORKLOADS[0])


def _pillar_world():
    def fn(x, y, z):
        out = np.where(y < 20, 1, 0)
        out = np.where((y == 19) & (x > 40), 4, out)  # a sand strip
        out = np.where((x >= 30) & (x < 34) & (z >= 30) & (z < 34) & (y >= 20) & (y < 40), 1, out)
        return np.where((x >= 10) & (x < 12) & (z >= 20) & (z < 22) & (y >= 20) & (y < 23), 11, out)  # glowstone
    w = World()
    w.insert_dense((0, 0, 0), H.dense_from(fn))
    return w


def _block_ids(world, v, vis):
    """Block id of each pixel's key (0 for a miss), through the oracle's decode."""
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
    dec = O.decode(sc, v.frame_bytes, vis)
    return np.where(vis.reshape(dec.shape) != 0, dec["material"], 0).reshape(v.frame.res[1], v.frame.res[0])


def _check(r, world, v, spec, hi, atlas, faces, tag, frames=1):
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
    hz = np.zeros(v.tiles[0] * v.tiles[1], np.uint32)
    r.set_view(v)
    for _ in range(frames):
        r.render(v.frame, spec, hi)
        vis, sun, hz, _ = O.render(sc, v.frame_bytes, O.spec_from(spec, hi), hz)
    r.resolve()
    r.sync()
    got = r.read_color()
    assert np.array_equal(r.read_vis(), vis), f"{tag}: keys"
    ref = O.resolve(sc, v.frame_bytes, O.rspec_from(spec, hi), atlas, SIZE, LAYERS, MIPS, faces, vis, sun)
    diff = got.view(np.uint32) != ref.view(np.uint32)
    assert not diff.any(), f"{tag}: {diff.any(axis=-1).sum()} pixels differ, first {np.argwhere(diff)[:3].tolist()}"
    return got


@pytest.mark.parametrize("spec_extra,hi_extra", [
    (0, 0), (0, LEGACY_LIGHTING), (SKY_TINT_BIT, SKY_SPECULAR), (LIGHT_RGB_BIT, EMITTER),
    (TEX_VARIATION_BIT | SKY_TINT_BIT, SKY_SPECULAR | SHORE_WET | SKY_COOL)])
def test_pillar_chunk_matches_oracle(spec_extra, hi_extra):
    world = _pillar_world()
    flags = A.FLAG_SHADOWS | 8 | A.FLAG_HIZ | spec_extra
    cam = H.camera((32.5, 40.3, 2.5), 90.0, -24.0, 160, 96, flags, sea=10.0)
    v = build_view(world, cam)
    v.frame.ambient, v.frame.mip_bias, v.frame.fog_density = 0.05, -5.0, 0.0
    v.frame.fog_scatter, v.frame.fog_g = 0.3, 0.6
    atlas, faces = _atlas(), _faces_perm()
    r = Renderer(160, 96)
    r.sync_world(world)
    r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
    got = _check(r, world, v, flags, A.FLAG_HI_FULL_MARCH | hi_extra, atlas, faces, f"pillar {spec_extra} {hi_extra}")
    assert (got[..., 3] == 1.0).mean() > 0.99
    r.close()


def test_small_world_views_match_oracle(small_world):
    """Random views of the 256x256x128 world: opaque pixels shaded, water and glass
    pixels left at alpha 0, sky elsewhere; fog with falloff and the sky look; a second
    frame with Hi-Z carried."""
    wl = W.WORKLOADS[0]
    flags = BASE_FLAGS | SKY_TINT_BIT
    cams = W.random_cameras(wl, 6, seed=23, resolutions=[(256, 144), (200, 120)], flags=flags)
    for c in cams[3:]:  # looking down onto the terrain and the sea
        c.pitch = -0.7
    atlas, faces = _atlas(7), _faces_perm()
    lane12 = 0
    for i, cam in enumerate(cams):
        v = build_view(small_world, cam)
        v.frame.ambient, v.frame.mip_bias = 0.04, -4.0
        v.frame.fog_density, v.frame.fog_falloff, v.frame.fog_height = 0.004, 0.02, 70.0
        v.frame.fog_scatter, v.frame.fog_g = 0.25, 0.7
        v.frame.haze_warm, v.frame.zenith_deep = 0.5, 0.4
        spec, hi = W.spec_words(cam.flags, small_world)
        hi = (hi | SKY_SPECULAR | SKY_LOOK) & ~(64 | 8192 | 16384 | 65536)
        r = Renderer(cam.width, cam.height)
        r.sync_world(small_world)
        r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
        print(f"resolve view {i}", flush=True)
        got = _check(r, small_world, v, spec, hi, atlas, faces, f"view {i}", frames=2)
        ids = _block_ids(small_world, v, r.read_vis())
        lane12 += int(((ids == 13) | (ids == 10)).sum())
        assert np.all(got[..., 3] == 1.0)  # lanes 1 and 2 shade the water and glass pixels
        r.close()
    assert lane12 > 0  # water or glass pixels were among them


def test_heatmap_frame_matches_oracle(small_world):
    wl = W.WORKLOADS[0]
    cam = W.default_camera(wl, flags=W.DEFAULT_FLAGS | A.FLAG_HEATMAP)
    v = build_view(small_world, cam)
    spec, hi = W.spec_words(cam.flags, small_world)
    sc = O.Scene(small_world.pool(0), small_world.pool(1), small_world.pool(2), v.chunks, v.grid)
    vis, sun, _, heat, _ = O.render_heat(sc, v.frame_bytes, O.spec_from(spec, hi),
                                         np.zeros(v.tiles[0] * v.tiles[1], np.uint32))
    atlas, faces = _atlas(3), _faces_perm()
    r = Renderer(cam.width, cam.height)
    r.sync_world(small_world)
    r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
    r.set_view(v)
    r.render(v.frame, spec, hi)
    r.resolve()
    r.sync()
    got = r.read_color()
    assert np.array_equal(r.read_heat(), heat)
    # a heatmap frame shades nothing: the distant-shadow and probe fields are not read
    ref = O.resolve(sc, v.frame_bytes, O.rspec_from(spec & ~(32768 | 67108864 | 134217728), hi), atlas, SIZE, LAYERS,
                    MIPS, faces, vis, sun, heat)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    r.close()


def test_unrestated_configurations_are_refused(small_world):
    wl = W.WORKLOADS[0]
    cam = W.default_camera(wl, flags=BASE_FLAGS)
    v = build_view(small_world, cam)
    spec, hi = W.spec_words(cam.flags, small_world)
    r = Renderer(cam.width, cam.height)
    r.sync_world(small_world)
    r.set_view(v)
    r.render(v.frame, spec, hi)
    with pytest.raises(RuntimeError):  # no assets yet
        r.resolve()
    r.set_resolve_assets(_atlas(), SIZE, LAYERS, MIPS, _faces())
    r.resolve()  # the base configuration is restated
    cases = [("foliage-rich albedo", lambda f: None, 0, 8192),
             ("canopy-relief albedo", lambda f: None, 0, 16384),
             ("cloud deck", lambda f: setattr(f, "cloud_cover", 0.4), 0, 0)]
    for name, frame_edit, spec_add, hi_add in cases:
        v2 = build_view(small_world, cam)
        frame_edit(v2.frame)
        r.render(v2.frame, spec | spec_add, hi | hi_add)
        with pytest.raises(RuntimeError, match="unsupported"):
            r.resolve()
        print(f"refused: {name}", flush=True)
    r.sync()
    r.close()


@pytest.mark.parametrize("encode,tone,grade,strength", [(1, 0, 0, 1.0), (0, 1, 1, 1.0), (1, 0, 2, 0.6)])
def test_present_matches_oracle(encode, tone, grade, strength):
    """The present blit of a resolved frame equals the oracle's bytes."""
    world = _pillar_world()
    flags = A.FLAG_SHADOWS | 8 | A.FLAG_HIZ | SKY_TINT_BIT
    cam = H.camera((32.5, 40.3, 2.5), 90.0, -24.0, 160, 96, flags, sea=10.0)
    v = build_view(world, cam)
    v.frame.ambient, v.frame.mip_bias, v.frame.fog_density = 0.05, -5.0, 0.003
    atlas, faces = _atlas(5), _faces_perm()
    r = Renderer(160, 96)
    r.sync_world(world)
    r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
    color = _check(r, world, v, flags, A.FLAG_HI_FULL_MARCH | SKY_SPECULAR, atlas, faces, "present")
    r.present(encode, tone, grade, strength)
    r.sync()
    got = r.read_image()
    want = O.present(color, encode, tone, grade, strength)
    assert np.array_equal(got, want), f"{(got != want).any(axis=-1).sum()} pixels differ"
    assert got[..., :3].std() > 5  # an image, not a flat fill
    r.close()


def test_taa_frames_match_oracle(small_world):
    """TAA (vx_taa) over four frames of a camera turning and moving: each frame's
    prev_view_proj is the previous camera's view_proj, the first frame's feedback is 1
    (no history), later ones 0.1, jittered from the second frame on. Every output half
    equals the oracle's TAA of the same resolve colour and history; the present after TAA
    draws the TAA output."""
    import copy
    import math

    from tests.test_taa import _vp
    from voxelcraft_menu_amd.world import CameraParams

    wl = W.WORKLOADS[0]
    flags = BASE_FLAGS | SKY_TINT_BIT
    atlas, faces = _atlas(3), _faces_perm()
    r = Renderer(200, 120)
    r.sync_world(small_world)
    r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
    base = W.default_camera(wl, flags=flags, width=200, height=120)
    prev_m = None
    hist = np.zeros((120, 200, 4), np.uint16)
    for k in range(4):
        cam = CameraParams(pos=(base.pos[0] + 0.7 * k, base.pos[1] - 0.3 * k, base.pos[2] + 0.2 * k),
                           yaw=base.yaw + 0.02 * k, pitch=base.pitch, width=200, height=120, sea_level=base.sea_level,
                           flags=flags)
        v = build_view(small_world, cam)
        fr = v.frame
        fr.ambient, fr.mip_bias = 0.04, -4.0
        if k:
            fr.jitter[0], fr.jitter[1] = (0.25, -0.125) if k % 2 else (-0.375, 0.3125)
        m = _vp(fr)
        fr.taa_feedback = 1.0 if k == 0 else 0.1
        pm = m if prev_m is None else prev_m
        for c in range(4):
            for rr in range(4):
                fr.prev_view_proj[c * 4 + rr] = float(pm[rr, c])
        prev_m = m
        spec, hi = W.spec_words(cam.flags, small_world)
        hi = (hi | SKY_SPECULAR) & ~(64 | 8192 | 16384 | 65536)
        r.set_view(v)
        r.render(fr, spec, hi)
        r.resolve()
        r.taa()
        r.sync()
        print(f"taa frame {k}", flush=True)
        sc = O.Scene(small_world.pool(0), small_world.pool(1), small_world.pool(2), v.chunks, v.grid)
        vis = r.read_vis()
        want = O.taa(sc, bytes(fr), vis, r.read_color(), hist)
        got = r.read_taa()
        diff = got != want
        assert not diff.any(), f"frame {k}: {diff.any(-1).sum()} pixels differ, first {np.argwhere(diff)[:3].tolist()}"
        if k:
            assert (got[..., :3] != O.taa(sc, bytes(copy.copy(fr)), vis, r.read_color(), np.zeros_like(hist))[..., :3]).any()
        hist = got
    r.present(encode=1)
    img = r.read_image()
    ref = O.present(hist.view(np.float16).astype(np.float32), encode=1)
    assert np.array_equal(img, ref)
    r.close()


@pytest.mark.parametrize("case", ["default", "per_pixel", "waves", "glass_reflect", "smooth_legs", "legacy"])
def test_water_and_glass_lanes_match_oracle(case):
    """Lanes 1 and 2 on the pool / land / glass-pane chunk of the closed-form pins, from
    three cameras: with SPEC_WATER_SEC and shore foam (the defaults), legs per pixel,
    waves on, SPEC_GLASS_REFLECT, smooth light at the leg hits, the legacy lighting and
    water; every colour bit-exact against the oracle."""
    from tests.test_oracle_lanes_pins import ABSORB, FLAT_SECONDARY, _dense
    world = World()
    world.insert_dense((0, 0, 0), _dense())
    flags = (A.FLAG_SHADOWS | 8 | A.FLAG_WATER_REFLECT | A.FLAG_WATER_REFRACT | A.FLAG_GLASS | A.FLAG_WATER_SHADOW_CUT
             | FLAT_SECONDARY | SKY_TINT_BIT)
    hi = A.FLAG_HI_FULL_MARCH | SKY_SPECULAR | 32 | 16
    if case == "per_pixel":
        hi &= ~32
    if case == "waves":
        flags |= 512 | 1024 | 2048 | 4096
    if case == "glass_reflect":
        hi |= A.FLAG_HI_GLASS_REFLECT
    if case == "smooth_legs":
        flags &= ~FLAT_SECONDARY
    if case == "legacy":
        hi |= LEGACY_LIGHTING | A.FLAG_HI_LEGACY_WATER
    atlas, faces = _atlas(11), _faces_perm()
    shaded = 0
    for k, (pos, yaw, pitch) in enumerate([((30.5, 56.3, 2.5), 78.0, -42.0), ((16.5, 50.3, 30.5), 90.0, -60.0),
                                           ((50.5, 44.3, 70.5), -100.0, -30.0)]):
        cam = H.camera(pos, yaw, pitch, 160, 96, flags, sea=26.0)
        v = build_view(world, cam)
        f = v.frame
        f.ambient, f.mip_bias, f.water_absorb = 0.05, -2.0, ABSORB
        f.fog_density, f.fog_falloff, f.fog_height = 0.003, 0.01, 30.0
        f.time, f.wave_amp, f.wave_scale, f.wave_speed, f.wave_reflect_slope = 2.5, 0.12, 32.0, 0.32, 0.35
        r = Renderer(160, 96)
        r.sync_world(world)
        r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
        print(f"lanes {case} camera {k}", flush=True)
        got = _check(r, world, v, flags, hi, atlas, faces, f"lanes {case} camera {k}")
        ids = _block_ids(world, v, r.read_vis())
        lanes = (ids == 13) | (ids == 10)
        assert np.all(got[..., 3][lanes] == 1.0)
        shaded += int(lanes.sum())
        r.close()
    assert shaded > 2000
