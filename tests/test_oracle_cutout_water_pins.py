"""The oracle's leaf-cutout micro march and water-interface probe against float64 brute force.

Leaf cutouts (leaf_solid, micro_entry, cutout_march; common.wgsl:514-577, called from
march_chunk at :858-879). A leaf voxel is cut into 4^3 micro cells; a micro cell is solid
when the lowbias32 hash of its world micro coordinate q = 4 * voxel + m, mixed as
(qx * 73856093) ^ (qy * 19349663) ^ (qz * 83492791), has (h >> 8) / 2^24 below the fill
(0.62 with FLAG_LEAF_THIN, else 0.72; src/render/mod.rs:321-323, :2500-2505). The
reference here is a float64 Amanatides-Woo DDA over the world at micro resolution (cells
of 1/4 voxel), where a leaf micro cell is solid iff (h >> 8) < fill * 2^24 in exact
integer terms and every other non-air voxel is solid in all its micro cells. The first
solid micro cell it meets gives the hit voxel, the micro cell, the face and t. The hash
is lowbias32 (pinned by its published inverse in test_oracle_pins.py and by the
constants of the reference's own tests/repeat.rs:186-194).

Water interface (common.wgsl:800-813). Without skip_water, a water voxel stops a ray
only when the cell the ray came from is not water; the probe reads that cell through the
chunk grid, so a ray crossing from a neighbouring chunk's water into the marched chunk's
water passes on. The reference is the float64 DDA through the dense world whose stop rule
is "non-air, and not (water entered from water)", restricted to the marched chunk's cells.

Each check counts disagreements. The shipped oracle must have none; copies of it built
with a planted mistake (VXO_PLANT in oracle/vxoracle.c) must have some, which shows the
checks can tell a wrong restatement from a right one.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from tests.test_oracle_frame_pins import LEAVES, STONE, WATER, _dda, _frame, _oracle, _rays, _stable, _world
from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd.world import World

PINE_LEAVES = 16
M32 = 0xFFFFFFFF
MIX = (73856093, 19349663, 83492791)


def _hash(x):
    """lowbias32 on uint32 arrays (common.wgsl:1422-1430)."""
    h = np.asarray(x, np.uint32).copy()
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x7FEB352D)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x846CA68B)
    h ^= h >> np.uint32(16)
    return h


def _unhash(x):
    """The published inverse of lowbias32 (Python ints)."""
    x ^= x >> 16
    x = (x * 0x43021123) & M32
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1D69E2A5) & M32
    x ^= x >> 16
    return x


def _mixed(q):
    qu = (np.asarray(q, np.int64) & M32).astype(np.uint32)
    return (qu[:, 0] * np.uint32(MIX[0])) ^ (qu[:, 1] * np.uint32(MIX[1])) ^ (qu[:, 2] * np.uint32(MIX[2]))


def _threshold(fill):
    # (h >> 8) * 2^-24 < f32(fill)  <=>  (h >> 8) < f32(fill) * 2^24, both sides exact
    return float(np.float32(fill)) * 2.0 ** 24


def leaf_micro_solid(q, fill):
    """Solid micro cells of leaf voxels; q = world micro coordinates, shape (n, 3)."""
    return (_hash(_mixed(q)) >> np.uint32(8)).astype(np.float64) < _threshold(fill)


class MicroSolid:
    """A dense block grid ([y, z, x], world voxel `base` at index 0) seen at micro
    resolution: `solid[y, z, x]` over micro cells, computed on demand."""

    def __init__(self, dense, fill, base=(0, 0, 0)):
        self.dense, self.fill, self.base = dense, fill, np.asarray(base, np.int64)
        self.shape = tuple(4 * s for s in dense.shape)

    def __getitem__(self, idx):
        y, z, x = (np.asarray(i, np.int64) for i in idx)
        ids = self.dense[y >> 2, z >> 2, x >> 2]
        solid = ids != 0
        leaf = (ids == LEAVES) | (ids == PINE_LEAVES)
        if leaf.any():
            q = np.stack([x[leaf], y[leaf], z[leaf]], 1) + 4 * self.base
            solid[leaf] = leaf_micro_solid(q, self.fill)
        return solid


def _key_micro(vis):
    v = vis.astype(np.uint64)
    return np.stack([(v >> np.uint64(22)) & np.uint64(3), (v >> np.uint64(14)) & np.uint64(3),
                     (v >> np.uint64(6)) & np.uint64(3)], 1).astype(np.int64)


def _fill_flags(fill):
    return A.FLAG_LEAF_CUTOUT | (A.FLAG_LEAF_THIN if fill == 0.62 else 0)


# ---- whole frames over a leaf canopy -----------------------------------------------------

def _canopy_frame_mismatches(fill, min_counts=True):
    """Disagreements of the oracle's frame (keys, micro cells, sun bits) with the micro DDA."""
    world, dense = _world(12, leaf_share=0.25)
    flags = A.FLAG_SHADOWS | _fill_flags(fill)
    cam, v = _frame(world, (60.5, 75.0, 40.5), 40.0, -20.0, flags=flags)
    spec = O.spec_from(flags, A.FLAG_HI_FULL_MARCH)
    assert spec.leaf_cutout == 1 and abs(spec.leaf_fill - fill) < 1e-6
    vis, sh, dec = _oracle(world, v, spec)
    ms = MicroSolid(dense, fill)
    rd = _rays(v)
    ro = np.tile(np.array(cam.pos, np.float64), (rd.shape[0], 1))
    (hit, cell, face, tm), ok = _stable(ms, 4.0 * ro, rd)
    ok &= ~hit | (face >= 0)
    m = ok & hit
    vox, micro, t = cell // 4, cell % 4, tm / 4.0
    leaf = np.zeros_like(m)
    leaf[m] = np.isin(dense[vox[m, 1], vox[m, 2], vox[m, 0]], (LEAVES, PINE_LEAVES))
    # leaf hits whose micro cell is not on the entry layer: the ray crossed empty micro cells first
    ax = np.where(m, face // 2, 0)
    entry_layer = np.where(rd[np.arange(rd.shape[0]), ax] > 0, 0, 3)
    deep = leaf & (micro[np.arange(rd.shape[0]), ax] != entry_layer)
    if min_counts:
        assert ok.mean() > 0.95, ok.mean()
        assert leaf.sum() > 500 and deep.sum() > 100, (leaf.sum(), deep.sum())
    bad = int(((vis != 0)[ok] != hit[ok]).sum())
    bad += int((~np.all(dec["voxel"][m] == vox[m], axis=1)).sum())
    bad += int((~np.all(_key_micro(vis[m]) == micro[m], axis=1)).sum())
    bad += int((dec["face"][m].astype(np.int64) != face[m]).sum())
    bad += int((np.abs(dec["t"][m] - t[m]) > 1e-4 * np.maximum(1.0, t[m])).sum())

    # sun rays where the guide is valid (dot(n, sun) > 0), from hit_t's point + 0.02 n
    sun = np.array(v.frame.sun_dir, np.float64)
    nvec = np.zeros((m.sum(), 3))
    a = face[m] // 2
    nvec[np.arange(a.size), a] = np.where(face[m] % 2 == 0, -1.0, 1.0)
    lit = nvec @ sun > 0
    idx = np.nonzero(m)[0][lit]
    o = ro[idx] + rd[idx] * dec["t"][idx].astype(np.float64)[:, None] + 0.02 * nvec[lit]
    (blk, bcell, _, _), ok2 = _stable(ms, 4.0 * o, np.tile(sun, (idx.size, 1)), max_t=4.0 * 220.0)
    if min_counts:
        leaf_block = blk & np.isin(dense[bcell[:, 1] // 4, bcell[:, 2] // 4, bcell[:, 0] // 4], (LEAVES, PINE_LEAVES))
        assert ok2.mean() > 0.95 and blk[ok2].sum() > 100 and (~blk[ok2]).sum() > 100
        assert (leaf_block & ok2).sum() > 20, "some sun rays must be stopped by a leaf micro cell"
    bad += int(((sh[idx] != 0)[ok2] != blk[ok2]).sum())
    bad += int((sh[np.nonzero(m)[0][~lit]] != 0).sum())
    return bad


@pytest.mark.parametrize("fill", [0.62, 0.72])
def test_canopy_frame_matches_micro_bruteforce(fill):
    assert _canopy_frame_mismatches(fill) == 0


# ---- march_chunk on random rays into a cutout chunk at negative coordinates --------------

def _leaf_chunk(key, seed):
    rng = np.random.default_rng(seed)
    y, z, x = np.meshgrid(np.arange(64), np.arange(64), np.arange(64), indexing="ij")
    dense = np.zeros((64, 64, 64), np.uint16)
    box = (x > 8) & (x < 56) & (y > 8) & (y < 56) & (z > 8) & (z < 56)
    r = rng.random(dense.shape)
    dense[box & (r < 0.30)] = LEAVES
    dense[box & (r >= 0.30) & (r < 0.36)] = PINE_LEAVES
    dense[box & (r >= 0.36) & (r < 0.37)] = STONE
    w = World()
    w.insert_dense(key, dense.reshape(-1))
    return w, dense


def _chunk_ray_mismatches(fill, key=(-3, 1, 5), n=2500, seed=5):
    w, dense = _leaf_chunk(key, seed)
    origin = np.array(key, np.int64) * 64
    scene = H.scene_for(w, H.gpu_chunk(w, key))
    frame = H.frame_for()
    spec = O.spec_from(_fill_flags(fill), A.FLAG_HI_FULL_MARCH)
    rng = np.random.default_rng(seed + 1)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ro_l = 32.0 + 70.0 * dirs                                  # chunk-local, outside the box
    tgt = rng.uniform(10.0, 54.0, size=(n, 3))
    rd = tgt - ro_l
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    ro32 = (ro_l + origin).astype(np.float32)
    rd32 = rd.astype(np.float32)
    got = [H.march_chunk(scene, frame, spec, 0, ro32[i], rd32[i]) for i in range(n)]
    ghit = np.array([g.hit for g in got], bool)
    gvox = np.array([list(g.voxel) for g in got], np.int64)
    gmic = np.array([list(g.micro) for g in got], np.int64)
    gface = np.array([g.normal for g in got], np.int64)
    gt = np.array([g.t for g in got], np.float64)
    ms = MicroSolid(dense, fill, base=origin)
    ro64 = ro32.astype(np.float64) - origin                    # the f32 origin, exactly
    (hit, cell, face, tm), ok = _stable(ms, 4.0 * ro64, rd32.astype(np.float64))
    ok &= ~hit | (face >= 0)
    # a march that gives up after MAX_ITERS = 512 iterations (common.wgsl:698-701) reports no
    # hit; a ray through many emptied leaf voxels can get there, and the DDA has no such cap
    capped = np.array([g.iters > 512 for g in got], bool)
    assert capped.sum() <= n // 500
    ok &= ~capped
    m = ok & hit
    leaf = np.zeros_like(m)
    leaf[m] = np.isin(dense[cell[m, 1] // 4, cell[m, 2] // 4, cell[m, 0] // 4], (LEAVES, PINE_LEAVES))
    assert ok.mean() > 0.95 and leaf.sum() > 0.5 * n
    bad = int((ghit[ok] != hit[ok]).sum())
    bad += int((~np.all(gvox[m] == cell[m] // 4, axis=1)).sum())
    bad += int((~np.all(gmic[m] == cell[m] % 4, axis=1)).sum())
    bad += int((gface[m] != face[m]).sum())
    bad += int((np.abs(gt[m] - tm[m] / 4.0) > 1e-4 * np.maximum(1.0, tm[m] / 4.0)).sum())
    return bad


@pytest.mark.parametrize("fill", [0.62, 0.72])
def test_cutout_march_chunk_matches_micro_bruteforce(fill):
    assert _chunk_ray_mismatches(fill) == 0


# ---- micro cells whose hash sits exactly on the threshold --------------------------------

def _cells_at(value, want=4):
    """World micro coordinates q with (hash(mix(q)) >> 8) == value, on a voxel face layer,
    found by inverting the hash and solving the mix for qx (the multipliers are odd)."""
    inv0 = pow(MIX[0], -1, 1 << 32)
    xs = np.array([_unhash((value << 8) | lo) for lo in range(256)], np.int64)
    qy, qz = np.meshgrid(np.arange(4, 240, dtype=np.int64), np.arange(-400, 400, dtype=np.int64), indexing="ij")
    qy, qz = qy.reshape(-1), qz.reshape(-1)
    rest = ((qy * MIX[1]) & M32) ^ ((qz * MIX[2]) & M32)
    out = []
    for x in xs:
        qx = (((x ^ rest) & M32) * inv0) & M32
        qx = np.where(qx >= 1 << 31, qx - (1 << 32), qx)
        for i in np.nonzero(np.abs(qx) < 1 << 13)[0]:
            q = np.array([qx[i], qy[i], qz[i]], np.int64)
            if np.any((q % 4 == 0) | (q % 4 == 3)):
                out.append(q)
        if len(out) >= want:
            break
    assert len(out) >= want
    return out[:want]


def _threshold_mismatches(fill):
    """One leaf voxel per chunk; a ray enters it through the face its chosen micro cell
    lies on. Cells with (h >> 8) == T are empty and cells with T - 1 solid, T = fill * 2^24."""
    T = int(_threshold(fill))
    assert T == _threshold(fill)
    spec = O.spec_from(_fill_flags(fill), A.FLAG_HI_FULL_MARCH)
    frame = H.frame_for()
    bad = 0
    for value, solid_cell in ((T, False), (T - 1, True)):
        for q in _cells_at(value):
            assert bool(leaf_micro_solid(q[None], fill)[0]) == solid_cell
            wv, mc = q // 4, q % 4
            key = tuple(int(k) for k in wv // 64)
            uc = wv % 64
            dense = np.zeros((64, 64, 64), np.uint16)
            dense[uc[1], uc[2], uc[0]] = LEAVES
            w = World()
            w.insert_dense(key, dense.reshape(-1))
            scene = H.scene_for(w, H.gpu_chunk(w, key))
            ax = int(np.nonzero((mc == 0) | (mc == 3))[0][0])
            sgn = 1.0 if mc[ax] == 0 else -1.0                 # enter through the face the cell lies on
            rd = np.array([0.013, -0.021, 0.017])
            rd[ax] = sgn
            rd /= np.linalg.norm(rd)
            centre = wv + (mc + 0.5) / 4.0
            ro = centre - rd * (3.0 + 0.125 / abs(rd[ax]))
            ro32, rd32 = ro.astype(np.float32), rd.astype(np.float32)
            g = H.march_chunk(scene, frame, spec, 0, ro32, rd32)
            base = np.array(key, np.int64) * 64
            ms = MicroSolid(dense, fill, base=base)
            (hit, cell, face, tm), ok = _stable(ms, 4.0 * (ro32.astype(np.float64)[None] - base),
                                                rd32.astype(np.float64)[None])
            assert ok[0]
            at_cell = bool(hit[0]) and np.array_equal(cell[0], (q - 4 * base))
            assert at_cell == solid_cell                      # the constructed case is what it claims
            bad += int(bool(g.hit) != bool(hit[0]))
            if hit[0] and g.hit:
                bad += int(list(g.voxel) != list(cell[0] // 4)) + int(list(g.micro) != list(cell[0] % 4))
                bad += int(g.normal != face[0]) + int(abs(g.t - tm[0] / 4.0) > 1e-4 * max(1.0, tm[0] / 4.0))
    return bad


@pytest.mark.parametrize("fill", [0.62, 0.72])
def test_threshold_cells_split_exactly_at_the_fill(fill):
    assert _threshold_mismatches(fill) == 0


# ---- the water-interface probe across a chunk border -------------------------------------

def _water_world():
    """Chunks A = (0,0,0) and B = (1,0,0): sea below y = 40 over a stone floor, stone
    pillars and an air pocket in B's water, a stone ledge in A."""
    rng = np.random.default_rng(21)
    y, z, x = np.meshgrid(np.arange(64), np.arange(64), np.arange(128), indexing="ij")
    dense = np.where(y < 40, WATER, 0)
    dense = np.where(y < 6, STONE, dense)
    for _ in range(10):
        px, pz, top = rng.integers(70, 124), rng.integers(2, 60), rng.integers(12, 46)
        dense[6:top, pz:pz + 3, px:px + 3] = STONE
    dense[18:30, 20:40, 84:96] = 0                              # air pocket inside B's water
    dense[20:24, 10:30, 30:40] = STONE                          # ledge in A
    dense = dense.astype(np.uint16)
    w = World()
    for cx in range(2):
        w.insert_dense((cx, 0, 0), np.ascontiguousarray(dense[:, :, 64 * cx:64 * (cx + 1)]).reshape(-1))
    return w, dense


def _water_interface_mismatches():
    w, dense = _water_world()
    chunks = np.concatenate([H.gpu_chunk(w, (0, 0, 0)), H.gpu_chunk(w, (1, 0, 0))])
    grid = np.full(A.VX_GRID_LEN, A.VX_NO_CHUNK, np.uint32)
    grid[0], grid[1] = 0, 1                                     # grid_min (0,0,0): x + 32 y + 256 z
    scene = H.scene_for(w, chunks, grid)
    frame = H.frame_for()
    spec = O.spec_from(A.FLAG_SHADOWS, A.FLAG_HI_FULL_MARCH)
    assert spec.water_repair == 1
    info = w.chunk_info((1, 0, 0))
    lo = np.array([64 + info.aabb_min[0], info.aabb_min[1], info.aabb_min[2]], np.int64)
    hi = np.array([64 + info.aabb_max[0], info.aabb_max[1], info.aabb_max[2]], np.int64)
    rng = np.random.default_rng(22)
    n = 1500
    wet = np.stack([rng.uniform(8, 60, n), rng.uniform(8, 38, n), rng.uniform(4, 60, n)], 1)
    dry = np.stack([rng.uniform(8, 60, n), rng.uniform(42, 62, n), rng.uniform(4, 60, n)], 1)
    ro = np.concatenate([wet, dry])
    tgt = np.stack([rng.uniform(64.5, 127.5, 2 * n), rng.uniform(1, 45, 2 * n), rng.uniform(0.5, 63.5, 2 * n)], 1)
    rd = tgt - ro
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    ro32, rd32 = ro.astype(np.float32), rd.astype(np.float32)
    got = [H.march_chunk(scene, frame, spec, 1, ro32[i], rd32[i]) for i in range(2 * n)]
    ghit = np.array([g.hit for g in got], bool)
    gvox = np.array([list(g.voxel) for g in got], np.int64) + np.array([64, 0, 0])
    gface = np.array([g.normal for g in got], np.int64)
    gt = np.array([g.t for g in got], np.float64)

    # ids of cells outside B's tight box carry +256: they never stop the ray but are its "previous cell"
    y, z, x = np.meshgrid(np.arange(64), np.arange(64), np.arange(128), indexing="ij")
    inside = (x >= lo[0]) & (x < hi[0]) & (y >= lo[1]) & (y < hi[1]) & (z >= lo[2]) & (z < hi[2])
    tagged = dense.astype(np.int64) + np.where(inside, 0, 256)

    def rule(cur, prev):
        return (cur < 256) & (cur != 0) & ((cur != WATER) | ((prev & 255) != WATER))

    (hit, vox, face, t), ok = _stable(tagged, ro32.astype(np.float64), rd32.astype(np.float64), rule=rule)
    ok &= ~hit | (face >= 0)
    m = ok & hit
    hid = np.zeros(2 * n, np.int64)
    hid[m] = dense[vox[m, 1], vox[m, 2], vox[m, 0]]
    passed = np.zeros(2 * n, bool)                              # rays that crossed water from A's water
    passed[:n] = m[:n] & (hid[:n] != WATER)
    assert ok.mean() > 0.95
    assert (hid[:n] == STONE).sum() > 300 and (hid[n:] == WATER).sum() > 300, "both sides of the rule occur"
    assert (m[:n] & (hid[:n] == WATER)).sum() > 10, "water entered from B's air pocket stops a ray from A's water"
    bad = int((ghit[ok] != hit[ok]).sum())
    bad += int((~np.all(gvox[m] == vox[m], axis=1)).sum())
    bad += int((gface[m] != face[m]).sum())
    bad += int((np.abs(gt[m] - t[m]) > 1e-4 * np.maximum(1.0, t[m])).sum())
    return bad


def test_water_interface_across_chunks_matches_bruteforce():
    assert _water_interface_mismatches() == 0


# ---- planted mistakes are caught ----------------------------------------------------------

@pytest.mark.parametrize("plant", [1, 2, 3, 4])
def test_planted_cutout_mistakes_are_caught(plant, tmp_path):
    with O.planted(plant, tmp_path):
        found = _threshold_mismatches(0.62)
        if not found:
            found = _chunk_ray_mismatches(0.62, n=600)
        if not found:
            found = _canopy_frame_mismatches(0.62, min_counts=False)
    assert found > 0, f"planted mistake {plant} went unnoticed"


@pytest.mark.parametrize("plant", [5, 6])
def test_planted_water_probe_mistakes_are_caught(plant, tmp_path):
    with O.planted(plant, tmp_path):
        assert _water_interface_mismatches() > 0, f"planted mistake {plant} went unnoticed"
