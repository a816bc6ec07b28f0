"""GPU parity of the distant terrain shadows -- the shaft scan (vx_shaft_scan) and the
resolve with FLAG_DISTANT_SHADOWS (terrain_shade_beyond on the sun term) -- and of the biome
tint, against the oracle through the C ABI, bit for bit.

- the scan of a height window of this build's generator, under two suns and two texel
  sizes, and of a spike on flat ground;
- resolved views of the 256x256x128 world under a low sun with the window around the
  camera (opaque pixels, and water and glass pixels whose legs' sun terms take the same
  factor), over two frames; the oracle without the factor differs on some pixels;
- a resolve with the flag before any upload reads the zeroed field;
- views with the workload's flags less probes (biome tint and distant shadows both on),
  with and without SPEC_TINT_BALANCE, at tint strengths 1 and 0.6;
- caustics over the shallow sea floor, and soft shadows (SPEC_SOFT_SHADOWS: the row and
  column passes and the mask read by lane 0) under a low sun, with distant shadows on in
  one view; the oracle's hard-shadow resolve differs on some pixels.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_resolve import BASE_FLAGS, LAYERS, MIPS, SIZE, SKY_SPECULAR, _atlas, _faces_perm
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.renderer import Renderer
from voxelcraft_menu_amd.world import ShaftField, build_view, set_tint

pytestmark = pytest.mark.gpu

DISTANT = 32768


def _frame(sun, texel):
    f = A.Frame()
    f.sun_dir[0], f.sun_dir[1], f.sun_dir[2] = sun
    f.shaft_texel = float(texel)
    return f


@pytest.fixture(scope="module")
def small_world():
    return W.make_world(W.WORKLOADS[0])


def test_scan_matches_oracle():
    r = Renderer(64, 64)
    for texel, cam, sun in ((4, (512.0, 300.0), (0.82, 0.04, 0.54)), (8, (-100.0, 2000.0), (-0.3, 0.03, 0.93))):
        sf = ShaftField(texel)
        sf.update(W.WORKLOADS[1].gen, *cam)
        h = sf.heights()
        fr = _frame(sun, texel)
        print(f"[shaft] scan texel {texel} sun {sun}", flush=True)
        r.upload_shaft(h)
        r.shaft_scan(fr)
        r.sync()
        got = r.read_shaft()
        want = O.shaft_scan(h, bytes(fr))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (texel, int((got != want).sum()))
        assert (got > h + 0.5).mean() > 0.01
        sf.close()
    spike = np.zeros((512, 512), np.float32)
    spike[256, 300] = 200.0
    fr = _frame((0.9396926, 0.34202015, 0.0), 4)
    r.upload_shaft(spike)
    r.shaft_scan(fr)
    r.sync()
    assert np.array_equal(r.read_shaft().view(np.uint32), O.shaft_scan(spike, bytes(fr)).view(np.uint32))
    r.close()


def test_resolve_with_distant_shadows_matches_oracle(small_world):
    wl = W.WORKLOADS[0]
    flags = BASE_FLAGS | DISTANT
    cams = W.random_cameras(wl, 3, seed=41, resolutions=[(256, 144), (200, 120)], flags=flags)
    cams[2].pitch = -0.6  # onto the terrain and the sea
    atlas, faces = _atlas(9), _faces_perm()
    changed = 0
    for i, cam in enumerate(cams):
        v = build_view(small_world, cam)
        v.frame.sun_dir[0], v.frame.sun_dir[1], v.frame.sun_dir[2] = 0.7962, 0.0839, 0.5991  # a low sun
        v.frame.ambient, v.frame.mip_bias = 0.04, -4.0
        v.frame.fog_density, v.frame.fog_falloff, v.frame.fog_height = 0.002, 0.02, 70.0
        sf = ShaftField()
        sf.update(wl.gen, v.frame.cam_pos[0], v.frame.cam_pos[2])
        sf.set_frame(v.frame)
        spec, hi = W.spec_words(cam.flags, small_world)
        hi = (hi | SKY_SPECULAR) & ~(64 | 8192 | 16384 | 65536)
        r = Renderer(cam.width, cam.height)
        r.sync_world(small_world)
        r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
        r.upload_shaft(sf.heights())
        r.shaft_scan(v.frame)
        r.set_view(v)
        sc = O.Scene(small_world.pool(0), small_world.pool(1), small_world.pool(2), v.chunks, v.grid)
        field = O.shaft_scan(sf.heights(), v.frame_bytes)
        hz = np.zeros(v.tiles[0] * v.tiles[1], np.uint32)
        print(f"[shaft] resolve view {i}", flush=True)
        for frame in range(2):
            r.render(v.frame, spec, hi)
            vis, sun, hz, _ = O.render(sc, v.frame_bytes, O.spec_from(spec, hi), hz)
            r.resolve()
            r.sync()
            got = r.read_color()
            assert np.array_equal(r.read_vis(), vis)
            ref = O.resolve(sc, v.frame_bytes, O.rspec_from(spec, hi, shaft=field), atlas, SIZE, LAYERS, MIPS, faces,
                            vis, sun)
            diff = got.view(np.uint32) != ref.view(np.uint32)
            assert not diff.any(), f"view {i} frame {frame}: {diff.any(axis=-1).sum()} pixels differ"
        plain = O.resolve(sc, v.frame_bytes, O.rspec_from(spec & ~DISTANT, hi), atlas, SIZE, LAYERS, MIPS, faces,
                          vis, sun)
        changed += int((plain.view(np.uint32) != ref.view(np.uint32)).any(axis=-1).sum())
        r.close()
        sf.close()
    assert changed > 0  # the distant term darkened some pixels


def test_resolve_before_any_upload_reads_the_zeroed_field(small_world):
    wl = W.WORKLOADS[0]
    cam = W.default_camera(wl, width=160, height=96, flags=BASE_FLAGS | DISTANT)
    v = build_view(small_world, cam)
    v.frame.shaft_texel, v.frame.shaft_soft = 4.0, 4.0
    v.frame.shaft_origin[0], v.frame.shaft_origin[1] = -1024.0, -1024.0
    spec, hi = W.spec_words(cam.flags, small_world)
    atlas, faces = _atlas(2), _faces_perm()
    r = Renderer(160, 96)
    r.sync_world(small_world)
    r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
    r.set_view(v)
    r.render(v.frame, spec, hi)
    r.resolve()
    r.sync()
    sc = O.Scene(small_world.pool(0), small_world.pool(1), small_world.pool(2), v.chunks, v.grid)
    vis, sun, _, _ = O.render(sc, v.frame_bytes, O.spec_from(spec, hi), np.zeros(v.tiles[0] * v.tiles[1], np.uint32))
    zero = np.zeros((512, 512), np.float32)
    ref = O.resolve(sc, v.frame_bytes, O.rspec_from(spec, hi, shaft=zero), atlas, SIZE, LAYERS, MIPS, faces, vis, sun)
    assert np.array_equal(r.read_color().view(np.uint32), ref.view(np.uint32))
    r.close()


def test_resolve_with_tint_and_distant_shadows_matches_oracle(small_world):
    wl = W.WORKLOADS[0]
    probes = 16777216 | 33554432 | 67108864 | 134217728 | 536870912 | 1073741824
    flags = W.DEFAULT_FLAGS & ~probes
    assert flags & 128 and flags & DISTANT
    cams = W.random_cameras(wl, 3, seed=43, resolutions=[(224, 128)], flags=flags)
    for c in cams[1:]:
        c.pitch = -0.5
    atlas, faces = _atlas(11), _faces_perm()
    for i, cam in enumerate(cams):
        v = build_view(small_world, cam)
        v.frame.ambient, v.frame.mip_bias = 0.05, -3.0
        set_tint(v.frame, 7 + i, (1.0, 0.6, 1.0)[i])
        sf = ShaftField()
        sf.update(wl.gen, v.frame.cam_pos[0], v.frame.cam_pos[2])
        sf.set_frame(v.frame)
        spec, hi = W.spec_words(cam.flags, small_world)
        hi = (hi | SKY_SPECULAR | (512 if i == 2 else 0)) & ~(64 | 8192 | 16384 | 65536)
        r = Renderer(cam.width, cam.height)
        r.sync_world(small_world)
        r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
        r.upload_shaft(sf.heights())
        r.shaft_scan(v.frame)
        r.set_view(v)
        r.render(v.frame, spec, hi)
        r.resolve()
        r.sync()
        print(f"[tint] view {i}", flush=True)
        sc = O.Scene(small_world.pool(0), small_world.pool(1), small_world.pool(2), v.chunks, v.grid)
        vis, sun, _, _ = O.render(sc, v.frame_bytes, O.spec_from(spec, hi),
                                  np.zeros(v.tiles[0] * v.tiles[1], np.uint32))
        field = O.shaft_scan(sf.heights(), v.frame_bytes)
        ref = O.resolve(sc, v.frame_bytes, O.rspec_from(spec, hi, shaft=field), atlas, SIZE, LAYERS, MIPS, faces, vis,
                        sun)
        got = r.read_color()
        diff = got.view(np.uint32) != ref.view(np.uint32)
        assert not diff.any(), f"view {i}: {diff.any(axis=-1).sum()} pixels differ"
        untinted = O.resolve(sc, v.frame_bytes, O.rspec_from(spec & ~128, hi, shaft=field), atlas, SIZE, LAYERS, MIPS,
                             faces, vis, sun)
        assert (untinted.view(np.uint32) != ref.view(np.uint32)).any()  # the tint changed the image
        r.close()
        sf.close()


def test_resolve_with_caustics_matches_oracle(small_world):
    """SPEC_CAUSTICS over the shallow sea floor of the 256x256x128 world, looking down."""
    wl = W.WORKLOADS[0]
    cams = W.random_cameras(wl, 3, seed=53, resolutions=[(224, 128)], flags=BASE_FLAGS)
    for c in cams:
        c.pitch = -0.9
    atlas, faces = _atlas(17), _faces_perm()
    changed = 0
    for i, cam in enumerate(cams):
        v = build_view(small_world, cam)
        v.frame.ambient, v.frame.mip_bias = 0.05, -3.0
        v.frame.wave_amp, v.frame.wave_scale, v.frame.wave_speed, v.frame.time = 0.4, 24.0, 0.3, 2.1
        spec, hi = W.spec_words(cam.flags, small_world)
        hi = (hi | SKY_SPECULAR | 64) & ~(8192 | 16384 | 65536)
        r = Renderer(cam.width, cam.height)
        r.sync_world(small_world)
        r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
        r.set_view(v)
        r.render(v.frame, spec, hi)
        r.resolve()
        r.sync()
        print(f"[caustics] view {i}", flush=True)
        sc = O.Scene(small_world.pool(0), small_world.pool(1), small_world.pool(2), v.chunks, v.grid)
        vis, sun, _, _ = O.render(sc, v.frame_bytes, O.spec_from(spec, hi),
                                  np.zeros(v.tiles[0] * v.tiles[1], np.uint32))
        ref = O.resolve(sc, v.frame_bytes, O.rspec_from(spec, hi), atlas, SIZE, LAYERS, MIPS, faces, vis, sun)
        got = r.read_color()
        diff = got.view(np.uint32) != ref.view(np.uint32)
        assert not diff.any(), f"view {i}: {diff.any(axis=-1).sum()} pixels differ"
        plain = O.resolve(sc, v.frame_bytes, O.rspec_from(spec, hi & ~64), atlas, SIZE, LAYERS, MIPS, faces, vis, sun)
        changed += int((plain.view(np.uint32) != ref.view(np.uint32)).any(axis=-1).sum())
        r.close()
    assert changed > 0  # caustics lit some sea-floor pixels


def test_resolve_with_soft_shadows_matches_oracle(small_world):
    """SPEC_SOFT_SHADOWS under a low sun (long blocker distances, penumbrae of several pixels)."""
    wl = W.WORKLOADS[0]
    cams = W.random_cameras(wl, 3, seed=59, resolutions=[(256, 144), (200, 120)], flags=BASE_FLAGS)
    for c in cams:
        c.pitch = -0.45
    cams[2].flags |= DISTANT
    atlas, faces = _atlas(19), _faces_perm()
    changed = 0
    for i, cam in enumerate(cams):
        v = build_view(small_world, cam)
        v.frame.sun_dir[0], v.frame.sun_dir[1], v.frame.sun_dir[2] = 0.7962, 0.0839, 0.5991
        v.frame.ambient, v.frame.mip_bias = 0.04, -3.0
        spec, hi = W.spec_words(cam.flags, small_world)
        hi = (hi | SKY_SPECULAR | 65536) & ~(64 | 8192 | 16384)
        r = Renderer(cam.width, cam.height)
        r.sync_world(small_world)
        r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
        sf = None
        if cam.flags & DISTANT:
            sf = ShaftField()
            sf.update(wl.gen, v.frame.cam_pos[0], v.frame.cam_pos[2])
            sf.set_frame(v.frame)
            r.upload_shaft(sf.heights())
            r.shaft_scan(v.frame)
        r.set_view(v)
        print(f"[soft] view {i} {cam.width}x{cam.height} distant {bool(cam.flags & DISTANT)}", flush=True)
        r.render(v.frame, spec, hi)
        r.resolve()
        r.sync()
        sc = O.Scene(small_world.pool(0), small_world.pool(1), small_world.pool(2), v.chunks, v.grid)
        vis, sun, _, _ = O.render(sc, v.frame_bytes, O.spec_from(spec, hi),
                                  np.zeros(v.tiles[0] * v.tiles[1], np.uint32))
        assert np.array_equal(r.read_vis(), vis)
        field = O.shaft_scan(sf.heights(), v.frame_bytes) if sf else None
        ref = O.resolve(sc, v.frame_bytes, O.rspec_from(spec, hi, shaft=field), atlas, SIZE, LAYERS, MIPS, faces,
                        vis, sun)
        got = r.read_color()
        diff = got.view(np.uint32) != ref.view(np.uint32)
        assert not diff.any(), f"view {i}: {diff.any(axis=-1).sum()} pixels differ"
        hard = O.resolve(sc, v.frame_bytes, O.rspec_from(spec, hi & ~65536, shaft=field), atlas, SIZE, LAYERS, MIPS,
                         faces, vis, sun)
        changed += int((hard.view(np.uint32) != ref.view(np.uint32)).any(axis=-1).sum())
        r.close()
        if sf:
            sf.close()
    assert changed > 0  # the penumbrae changed some pixels
