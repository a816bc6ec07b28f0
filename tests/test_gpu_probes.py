"""GPU parity of the ambient probe field: its layout (vx_probe_fill / vx_probe_write /
vx_read_probe against make_probe and write_probe's placement, restated here in numpy) and
the resolve with the probe cube, bounce and sun (FLAG_PROBE_CUBE | _BOUNCE | _SUN, the
workload's defaults), bit for bit against the oracle over the same field.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_resolve import LAYERS, MIPS, SIZE, SKY_SPECULAR, _atlas, _faces_perm
from voxelcraft_menu_amd import workloads as W
from voxelcraft_menu_amd.renderer import Renderer
from voxelcraft_menu_amd.world import ShaftField, build_view, probe_bake, set_tint

pytestmark = pytest.mark.gpu

CUBE, BOUNCE, SUN = 67108864, 134217728, 536870912


def _ones():
    return np.full((128, 768, 128, 4), 0x3C00, np.uint16)


def _write(field, origin, occl, bounce):
    """write_probe's placement (src/render/mod.rs:1398-1445) of one chunk's bake."""
    bx, bz = (origin[0] // 4) % 128, (origin[2] // 4) % 128
    by = min(max(origin[1] // 4, 0), 112)
    f = field.view(np.float16)
    for axis in range(6):
        blk = f[bz:bz + 16, axis * 128 + by:axis * 128 + by + 16, bx:bx + 16]
        blk[..., :3] = bounce[axis].astype(np.float16)
        blk[..., 3] = occl[axis].astype(np.float16)


def _noise():
    i = np.arange(128 * 768 * 128, dtype=np.uint64)
    out = np.empty((i.size, 4), np.float16)
    for c in range(4):
        h = ((i * 2654435761 + c * 0x9E3779B9) & 0xFFFFFFFF).astype(np.uint32)
        out[:, c] = (np.float32(0.5) + (h >> 8).astype(np.float32) / np.float32(1 << 25)).astype(np.float16)
    return out.view(np.uint16).reshape(128, 768, 128, 4)


def test_probe_field_layout():
    gen = W.WORKLOADS[1].gen
    r = Renderer(64, 64)
    assert np.array_equal(r.read_probe(), _ones())  # a context starts with the field of ones
    want = _ones()
    for origin in ((0, 64, 0), (-64, 128, 448), (960, 0, -512), (64, 192, 64)):
        print(f"[probe] write {origin}", flush=True)
        occl, bounce = probe_bake(gen, origin)
        r.write_probe(origin, occl, bounce)
        _write(want, origin, occl, bounce)
    assert np.array_equal(r.read_probe(), want)
    r.fill_probe(noise=True)
    assert np.array_equal(r.read_probe(), _noise())
    r.fill_probe(0.25)  # pinned: later writes are ignored
    occl, bounce = probe_bake(gen, (0, 64, 0))
    r.write_probe((0, 64, 0), occl, bounce)
    assert np.all(r.read_probe() == np.float16(0.25).view(np.uint16))
    r.close()


@pytest.mark.parametrize("extra", [0, 16777216, 33554432], ids=["workload", "probe-tap", "probe-ambient"])
def test_resolve_with_probes_matches_oracle(extra):
    """The workload's flags, then with SPEC_PROBE_TAP and with SPEC_PROBE_AMBIENT added."""
    wl = W.WORKLOADS[0]
    world = W.make_world(wl)
    assert W.DEFAULT_FLAGS & CUBE and W.DEFAULT_FLAGS & BOUNCE and W.DEFAULT_FLAGS & SUN
    assert not W.DEFAULT_FLAGS & (16777216 | 33554432)
    flags = W.DEFAULT_FLAGS | extra
    cams = W.random_cameras(wl, 3, seed=47, resolutions=[(224, 128)], flags=flags)
    for c in cams[1:]:
        c.pitch = -0.5
    atlas, faces = _atlas(13), _faces_perm()
    field = _ones()
    r = Renderer(224, 128)
    r.sync_world(world)
    r.set_resolve_assets(atlas, SIZE, LAYERS, MIPS, faces)
    for key in world.keys():
        origin = tuple(int(k) * 64 for k in key)
        occl, bounce = probe_bake(wl.gen, origin)
        r.write_probe(origin, occl, bounce)
        _write(field, origin, occl, bounce)
    assert np.array_equal(r.read_probe(), field)
    for i, cam in enumerate(cams):
        v = build_view(world, cam)
        v.frame.ambient, v.frame.mip_bias = 0.06, -3.0
        set_tint(v.frame, wl.gen.seed)
        sf = ShaftField()
        sf.update(wl.gen, v.frame.cam_pos[0], v.frame.cam_pos[2])
        sf.set_frame(v.frame)
        spec, hi = W.spec_words(cam.flags, world)
        hi = (hi | SKY_SPECULAR) & ~(64 | 8192 | 16384 | 65536)
        r.upload_shaft(sf.heights())
        r.shaft_scan(v.frame)
        r.set_view(v)
        r.render(v.frame, spec, hi)
        r.resolve()
        r.sync()
        print(f"[probe] resolve view {i} extra {extra}", flush=True)
        sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
        vis, sun, _, _ = O.render(sc, v.frame_bytes, O.spec_from(spec, hi),
                                  np.zeros(v.tiles[0] * v.tiles[1], np.uint32))
        shaft = O.shaft_scan(sf.heights(), v.frame_bytes)
        ref = O.resolve(sc, v.frame_bytes, O.rspec_from(spec, hi, shaft=shaft, probe=field), atlas, SIZE, LAYERS,
                        MIPS, faces, vis, sun)
        got = r.read_color()
        diff = got.view(np.uint32) != ref.view(np.uint32)
        assert not diff.any(), f"view {i}: {diff.any(axis=-1).sum()} pixels differ"
        plain = O.resolve(sc, v.frame_bytes, O.rspec_from(spec & ~(CUBE | BOUNCE | SUN | 16777216 | 33554432), hi,
                                                          shaft=shaft), atlas,
                          SIZE, LAYERS, MIPS, faces, vis, sun)
        assert (plain.view(np.uint32) != ref.view(np.uint32)).any()  # the probes changed the image
        sf.close()
    r.close()
