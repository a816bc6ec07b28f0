"""Device chunk build (build_local on the GPU, src/voxel/tree.rs:89-204).

The host's build_local is pinned by the tree tests (test_oracle_pins.py, after
tests/tree.rs); here its flat record is checked to carry the same tree (a world built
from the records equals one built from the dense chunks, pool byte for pool byte), and
the device build must write the same record bytes for every chunk.
"""
import numpy as np
import pytest

from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.world import World, build_local_flat

STONE, DIRT, GLASS, GLOWSTONE, WATER, LEAVES, TALL_GRASS, REEDS = 1, 2, 10, 11, 13, 7, 18, 25


def _grid():
    y, z, x = np.meshgrid(np.arange(64), np.arange(64), np.arange(64), indexing="ij")
    return x, y, z


def chunk_cases(seed=5):
    """Dense chunks (x | z << 6 | y << 12 order) covering build_local's branches."""
    rng = np.random.default_rng(seed)
    x, y, z = _grid()
    cases = {}
    cases["air"] = np.zeros((64, 64, 64), np.uint16)
    cases["stone"] = np.full((64, 64, 64), STONE, np.uint16)
    cases["two_ids_full"] = np.where((x + y + z) % 2 == 0, STONE, DIRT).astype(np.uint16)
    cases["full_l1_blocks"] = np.where(((x // 16) + (y // 16) + (z // 16)) % 2 == 0, STONE, 0).astype(np.uint16)
    cases["water_layers"] = np.where(y < 20, STONE, np.where(y < 41, WATER, 0)).astype(np.uint16)
    cases["water_only"] = np.where(y < 30, WATER, 0).astype(np.uint16)
    many = rng.integers(0, 26, size=(64, 64, 64)).astype(np.uint16)  # up to 25 ids per leaf: 8-bit indices
    cases["many_ids"] = many
    cases["ids_past_table"] = np.where(rng.random((64, 64, 64)) < 0.5, rng.integers(26, 400, (64, 64, 64)),
                                       0).astype(np.uint16)
    sparse = np.zeros((64, 64, 64), np.uint16)
    pick = rng.random((64, 64, 64)) < 0.01
    sparse[pick] = rng.choice([STONE, GLASS, GLOWSTONE, LEAVES, TALL_GRASS, REEDS, WATER], size=int(pick.sum()))
    cases["sparse_mixed"] = sparse
    few = np.where(y < 32 + (x % 7), rng.choice([STONE, DIRT, GLASS, GLOWSTONE], size=(64, 64, 64)), 0)
    cases["terrain_like"] = few.astype(np.uint16)
    for k in (3, 5):  # palettes of 3 / 5 ids (2- and 4-bit indices), odd counts
        cases[f"ids_{k}"] = np.where(z > 10, rng.integers(1, k + 1, size=(64, 64, 64)), 0).astype(np.uint16)
    corner = np.zeros((64, 64, 64), np.uint16)
    corner[63, 63, 63] = GLOWSTONE
    corner[0, 0, 0] = WATER
    cases["two_corners"] = corner
    return {k: v.reshape(-1) for k, v in cases.items()}


def _header(rec: bytes) -> A.LocalTree:
    return A.LocalTree.from_buffer_copy(rec[:64])


def test_flat_record_is_the_host_tree():
    cases = chunk_cases()
    a, b = World(), World()
    for i, (name, dense) in enumerate(cases.items()):
        key = (i % 4, i // 4 % 4, i // 16)
        a.insert_dense(key, dense)
        rec = build_local_flat(dense)
        h = _header(rec)
        assert h.bytes == len(rec) and len(rec) % 8 == 0, name
        assert h.n_l1 == bin(h.root_mask).count("1"), name
        assert h.solid_count == int((dense != 0).sum()), name
        b.insert_local_flat(key, rec)
    for pool in range(3):
        assert a.pool(pool).tobytes() == b.pool(pool).tobytes(), pool


def test_flat_record_fields():
    cases = chunk_cases()
    h = _header(build_local_flat(cases["stone"]))
    assert h.flags & 1 and h.uniform == STONE and h.n_attrs == 0 and h.n_palette == 0  # uniform chunk
    assert h.n_leaves == 0 and h.n_l1 == 64  # full L1 cells drop their leaves
    h = _header(build_local_flat(cases["air"]))
    assert h.root_mask == 0 and h.solid_count == 0 and bytes(h.aabb_max) == b"\0\0\0"
    h = _header(build_local_flat(cases["water_only"]))
    assert h.flags & 2 and h.dry_mask == 0  # water but nothing dry
    h = _header(build_local_flat(cases["two_corners"]))
    assert bytes(h.aabb_min) == b"\0\0\0" and bytes(h.aabb_max) == b"\x40\x40\x40" and h.emitters == 1


@pytest.mark.gpu
def test_device_build_matches_host_records():
    from voxelcraft_menu_amd import workloads as W
    from voxelcraft_menu_amd.renderer import Renderer

    cases = chunk_cases()
    names = list(cases)
    dense = np.stack([cases[k] for k in names])
    # chunks of a generated world (terrain, sea, caves, trees)
    wl = W.WORKLOADS[0]
    world = W.make_world(wl)
    keys = [tuple(int(v) for v in k) for k in world.keys()][:40]
    gen = np.stack([world.extract_dense(k) for k in keys])
    # more than one device pass (256 chunks per pass)
    rng = np.random.default_rng(11)
    extra = np.stack([np.where(rng.random(64 ** 3) < p, rng.integers(1, 26, 64 ** 3), 0).astype(np.uint16)
                      for p in np.linspace(0.0, 1.0, 230)])
    batch = np.concatenate([dense, gen, extra])
    print(f"device build of {batch.shape[0]} chunks", flush=True)
    r = Renderer(64, 64)
    recs = r.build_local(batch)
    r.close()
    for i in range(batch.shape[0]):
        want = build_local_flat(batch[i])
        label = names[i] if i < len(names) else f"chunk {i}"
        assert recs[i] == want, label
