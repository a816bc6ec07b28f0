"""The C ABI: every declared symbol is exported, struct layouts match the reference,
and host-side argument checks behave without a GPU."""
import ctypes as C
import os
import re

import pytest

from voxelcraft_menu_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b((?:vx|vxw)_[a-z0-9_]+)\s*\(", src)))


@pytest.mark.parametrize("header,loader", [("vxhost.h", A.host), ("vxmarch.h", A.march)])
def test_every_declared_symbol_is_exported(header, loader):
    lib = loader()
    names = _declared(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_headers():
    assert set(_declared("vxhost.h")) <= set(A.host_symbols())
    assert set(_declared("vxmarch.h")) <= set(A.march_symbols())


def test_struct_layouts_match_the_reference():
    # GpuFrame offsets asserted in src/render/mod.rs:458-466 and validation/check_compact.py:45
    assert C.sizeof(A.Frame) == 352
    assert A.Frame.cloud_cover.offset == 192 and A.Frame.wave_amp.offset == 240
    assert A.Frame.shaft_origin.offset == 256 and A.Frame.zenith_deep.offset == 284
    assert A.Frame.prev_view_proj.offset == 288
    assert A.Frame.grid_min.offset == 144 and A.Frame.sea_level.offset == 156
    assert A.Frame.grid_dim.offset == 160 and A.Frame.jitter.offset == 176 and A.Frame.frame_index.offset == 188
    # GpuChunk (src/render/mod.rs:468-489)
    assert C.sizeof(A.Chunk) == 96
    assert A.Chunk.root.offset == 48 and A.Chunk.attr_base.offset == 64 and A.Chunk.dry_mask.offset == 80
    assert C.sizeof(A.Edit) == 16


def test_error_strings_and_no_device_path():
    L = A.march()
    assert L.vx_error_string(0) == b"ok"
    assert L.vx_error_string(-5).startswith(b"feature")
    # argument checks run before any device work
    assert L.vx_ctx_create(None, None) == -3
    assert L.vx_render(None, None, 0, 0) == -3
    assert L.vx_pool_sync(None, 0, None, 0, None, 0) == -3
    out = C.c_void_p()
    assert L.vx_batch_create(None, 4, C.byref(out)) == -3
    assert L.vx_batch_set_views(None, None, 1) == -3
    assert L.vx_batch_render(None) == -3
    assert L.vx_batch_set_tuning(None, None) == -3
    assert L.vx_batch_autotune(None, 3, None, None) == -3
    assert L.vx_batch_read_vis(None, 0, None, 0) == -3
    assert L.vx_serve_create(None, None, C.byref(out)) == -3
    assert L.vx_serve_submit(None, 0, None, 0.0) == -3
    assert L.vx_serve_release(None, 0) == -3
    assert L.vx_serve_barrier(None) == -3
    assert L.vx_serve_clock() > 0.0
    if os.environ.get("HIP_VISIBLE_DEVICES") == "" or not os.path.exists("/dev/kfd"):
        ctx = C.c_void_p()
        desc = A.CtxDesc(64, 64, 0, 0, -1)
        assert L.vx_ctx_create(C.byref(desc), C.byref(ctx)) == -1  # VX_E_NODEV


def test_host_argument_checks():
    L = A.host()
    assert L.vxw_insert_dense(None, 0, 0, 0, None) == -3
    assert L.vxw_build_view(None, None, None, None, None) == -3
    assert L.vxw_pack_frame(None, None, 0, None) == -3


def test_struct_sizes_match_the_c_header(tmp_path):
    """ctypes mirrors against the C compiler's view of include/vxmarch.h."""
    import subprocess

    names = {"vx_gpu_frame": A.Frame, "vx_gpu_chunk": A.Chunk, "vx_edit": A.Edit, "vx_range": A.Range,
             "vx_ctx_desc": A.CtxDesc, "vx_tuning": A.Tuning, "vx_pair_demand": A.PairDemand,
             "vx_timings": A.Timings, "vx_hit": A.Hit, "vx_view": A.ViewDesc, "vx_camera_view": A.CameraView, "vx_batch_tuning": A.BatchTuning,
             "vx_serve_desc": A.ServeDesc, "vx_serve_result": A.ServeResult, "vx_serve_stats": A.ServeStats,
             "vx_march_stats": A.MarchStats, "vx_pool_mirror": A.PoolMirror}
    names_np = {"vx_sec_hit": A.SEC_DTYPE}
    src = tmp_path / "sz.c"
    body = "".join(f'    printf("{n} %zu\\n", sizeof({n}));\n' for n in list(names) + list(names_np))
    src.write_text(f'#include <stdio.h>\n#include "vxmarch.h"\nint main(void) {{\n{body}    return 0;\n}}\n')
    exe = tmp_path / "sz"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c99", f"-I{inc}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    got = dict(zip(out[0::2], (int(x) for x in out[1::2])))
    for n, cls in names.items():
        assert got[n] == C.sizeof(cls), (n, got[n], C.sizeof(cls))
    for n, dt in names_np.items():
        assert got[n] == dt.itemsize, (n, got[n], dt.itemsize)


def test_host_struct_sizes_match_the_c_header(tmp_path):
    """ctypes / numpy mirrors against the C compiler's view of include/vxhost.h."""
    import subprocess

    names = {"vxw_chunk_info": A.ChunkInfo, "vxw_key": A.Key, "vxw_render_item": A.RenderItem,
             "vxw_lod_config": A.LodConfig, "vxw_camera": A.Camera, "vxw_view": A.View,
             "vxw_gen_params": A.GenParams, "vxw_player": A.Player, "vxw_journal_info": A.JournalInfo}
    names_np = {"vxw_render_item": A.RENDER_ITEM_DTYPE, "vxw_key": A.KEY_DTYPE}
    src = tmp_path / "szh.c"
    body = "".join(f'    printf("{n} %zu\\n", sizeof({n}));\n' for n in names)
    src.write_text(f'#include <stdio.h>\n#include "vxhost.h"\nint main(void) {{\n{body}    return 0;\n}}\n')
    exe = tmp_path / "szh"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c99", f"-I{inc}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    got = dict(zip(out[0::2], (int(x) for x in out[1::2])))
    for n, cls in names.items():
        assert got[n] == C.sizeof(cls), (n, got[n], C.sizeof(cls))
    for n, dt in names_np.items():
        assert got[n] == dt.itemsize, (n, got[n], dt.itemsize)


def test_noalloc_checks_its_arguments_and_tile_limit_on_the_host():
    """vx_render_noalloc refuses missing buffers and frames beyond VX_NOALLOC_MAX_TILES
    before it touches a device (the pointers below are never dereferenced)."""
    L = A.march()
    assert C.sizeof(A.DeviceView) == 72 and A.DeviceView.grid.offset == 64
    dv = A.DeviceView(inners=8, inners_bytes=16, leaves=8, leaves_bytes=8, bricks=8, bricks_bytes=4,
                      chunks=8, n_chunks=1, grid=8)
    f = A.Frame()
    f.res[0], f.res[1], f.tiles[0], f.tiles[1] = 4096, 4096, 512, 512
    f.grid_dim[0], f.grid_dim[1], f.grid_dim[2] = 32, 8, 32
    buf = C.c_void_p(8)
    assert 512 * 512 > A.NOALLOC_MAX_TILES
    assert L.vx_render_noalloc(C.byref(f), 0, 0, C.byref(dv), buf, buf, buf, None) == -4  # VX_E_OVERFLOW
    assert L.vx_render_noalloc(C.byref(f), 0, 0, C.byref(dv), None, buf, buf, None) == -3
    assert L.vx_render_noalloc(C.byref(f), 0, 0, None, buf, buf, buf, None) == -3
    f.flags = A.FLAG_HEATMAP
    f.res[0], f.res[1], f.tiles[0], f.tiles[1] = 64, 64, 8, 8
    assert L.vx_render_noalloc(C.byref(f), 0, 0, C.byref(dv), buf, buf, buf, None) == -5


def test_a_library_built_from_other_sources_is_refused(monkeypatch):
    from voxelcraft_menu_amd import build as B
    A.check_fresh(A.MARCH_LIB, "march")
    A.check_fresh(A.HOST_LIB, "host")
    monkeypatch.setattr(B, "source_digest", lambda kind: "0" * 64)
    with pytest.raises(RuntimeError, match="does not match its sources"):
        A.check_fresh(A.MARCH_LIB, "march")
