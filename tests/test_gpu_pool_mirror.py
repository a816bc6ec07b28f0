"""vx_pool_sync keeps GpuMirror::sync's rules (src/render/pools.rs:44-98), checked on the
device mirror itself (vx_pool_read) and on its upload counter, with byte counts worked
out by hand for each case:

- dirty ranges (element offset, length) become byte ranges, sorted, and a range starting
  at most 4096 bytes after the previous one's end merges with it (the gap is uploaded
  too);
- more than 512 merged ranges upload the whole pool;
- ranges are clipped to the pool's length, ends rounded up to 4 bytes;
- a pool longer than the mirror's capacity doubles the capacity (from 1 MiB) and uploads
  the whole pool.
The first sync of a context uploads the whole pool (a deviation: a new mirror has no
contents, and another mirror may already have drained the dirty list).
"""
import ctypes as C

import numpy as np
import pytest

from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.renderer import Renderer

pytestmark = pytest.mark.gpu
MIB = 1 << 20


def _mirror(r, pool):
    m = A.PoolMirror()
    A.check(A.march().vx_pool_mirror_get(r.handle, pool, C.byref(m)), "vx_pool_mirror_get")
    return m


def _read(r, pool, n):
    out = np.empty(n, np.uint8)
    A.check(A.march().vx_pool_read(r.handle, pool, A.ptr(out), n), "vx_pool_read")
    return out


def _sync(r, pool, data, dirty):
    rng = np.array(dirty, dtype=np.uint32).reshape(-1, 2)
    A.check(A.march().vx_pool_sync(r.handle, pool, A.ptr(data), data.nbytes, A.ptr(rng) if len(rng) else None,
                                   len(rng)), "vx_pool_sync")


def test_dirty_ranges_merge_clip_and_round_as_the_reference():
    rng = np.random.default_rng(5)
    r = Renderer(64, 64)
    P = A.POOL_LEAVES  # 8-byte elements
    d0 = rng.integers(0, 256, 65536, dtype=np.uint8)
    _sync(r, P, d0, [])
    m = _mirror(r, P)
    assert (m.cap, m.len, m.uploaded_bytes) == (MIB, 65536, 65536)
    assert np.array_equal(_read(r, P, 65536), d0)

    # [0, 16) and [4112, 4120): 4112 <= 16 + 4096, one range [0, 4120)
    d1 = d0 ^ np.uint8(0x5A)
    _sync(r, P, d1, [(0, 2), (514, 1)])
    got = _read(r, P, 65536)
    assert np.array_equal(got[:4120], d1[:4120]) and np.array_equal(got[4120:], d0[4120:])
    assert _mirror(r, P).uploaded_bytes == 65536 + 4120

    # [0, 16) and [4120, 4128): 4120 > 16 + 4096, two ranges of 16 and 8 bytes; unsorted input
    d2 = d1 ^ np.uint8(0xA5)
    _sync(r, P, d2, [(515, 1), (0, 2)])
    got = _read(r, P, 65536)
    assert np.array_equal(got[:16], d2[:16]) and np.array_equal(got[16:4120], d1[16:4120])
    assert np.array_equal(got[4120:4128], d2[4120:4128]) and np.array_equal(got[4128:], d0[4128:])
    assert _mirror(r, P).uploaded_bytes == 65536 + 4120 + 24

    # past the end: nothing; across the end: clipped to [65520, 65536)
    _sync(r, P, d2, [(9000, 4)])
    assert _mirror(r, P).uploaded_bytes == 65536 + 4120 + 24
    _sync(r, P, d2, [(8190, 10)])
    assert _mirror(r, P).uploaded_bytes == 65536 + 4120 + 24 + 16
    assert np.array_equal(_read(r, P, 65536)[65520:], d2[65520:])
    r.close()


def test_more_than_512_ranges_upload_the_pool_and_growth_doubles():
    rng = np.random.default_rng(6)
    r = Renderer(64, 64)
    P = A.POOL_BRICKS  # 4-byte elements
    n = 4 * MIB
    d0 = rng.integers(0, 256, n, dtype=np.uint8)
    _sync(r, P, d0, [])  # longer than the 1 MiB mirror: capacity 1 -> 2 -> 4 MiB, whole upload
    m0 = _mirror(r, P)
    assert (m0.cap, m0.len, m0.uploaded_bytes) == (4 * MIB, n, n)
    assert np.array_equal(_read(r, P, n), d0)

    # 512 one-element ranges 8 KiB apart: 512 ranges of 4 bytes
    d1 = d0 ^ np.uint8(0x33)
    _sync(r, P, d1, [(i * 2048, 1) for i in range(512)])
    got = _read(r, P, n)
    idx = (np.arange(512) * 8192)[:, None] + np.arange(4)
    mask = np.zeros(n, bool)
    mask[idx.ravel()] = True
    assert np.array_equal(got[mask], d1[mask]) and np.array_equal(got[~mask], d0[~mask])
    assert _mirror(r, P).uploaded_bytes == n + 512 * 4

    # 513 such ranges: the whole pool
    _sync(r, P, d1, [(i * 2048, 1) for i in range(513)])
    assert np.array_equal(_read(r, P, n), d1)
    assert _mirror(r, P).uploaded_bytes == n + 512 * 4 + n

    # growth past the capacity: 4 -> 8 MiB, whole upload, the pool generation moves
    d2 = rng.integers(0, 256, 5 * MIB, dtype=np.uint8)
    _sync(r, P, d2, [(0, 1)])
    m2 = _mirror(r, P)
    assert (m2.cap, m2.len, m2.uploaded_bytes) == (8 * MIB, 5 * MIB, n + 2048 + n + 5 * MIB)
    assert m2.generation != m0.generation
    assert np.array_equal(_read(r, P, 5 * MIB), d2)
    r.close()
