"""The biome tint (resolve.wgsl:166-287) in the oracle, pinned by what the noise and the
tint table fix.

biome_noise is 2D OpenSimplex2 (FastNoiseLite 1.1.1's SingleOpenSimplex2, which
tests/biome.rs cross-checks against the crate; the crate is not in this container), so:
- it is 0 at every vertex of the skewed lattice (the vertex's own kernel sees a zero
  offset, the others are beyond their radius);
- near a vertex it is linear with gradient 99.83685 x 0.5^4 x g = 6.2398 g in noise units
  (freq x world), g a unit vector at 7.5 + 15 k degrees (the 24- and 8-direction tables),
  the same from every side: one-sided slopes in eight directions around the vertex agree
  with the central-difference gradient;
- its extrema over a wide sample approach the normalisation bound 1 and never pass it.
biome_tint is the humidity-weighted mix of the dry / mid / wet rows of the table, each the
temperature-weighted mix of its biomes, with band weights (cold, 1 - cold - hot, hot)
from smoothsteps at +-0.22 +- 0.18: checked in float64 from the oracle's two noise values.
Through the resolve, on the floor-and-pillar chunk of the lane-0 pins: tint strength 0 is
bit-identical to no tint, and each floor pixel is the lane-0 closed form with the albedo
times mix(1, tint(hit.xz), alpha x strength), alpha the atlas texel's (128 / 255 here).
Planted mistakes 53-56 each fail one of these.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import test_oracle_resolve_pins as RP
from voxelcraft_menu_amd.world import TINT_HUMID_FREQ, TINT_TEMP_FREQ, set_tint

G2 = (3.0 - math.sqrt(3.0)) / 6.0
KERNEL = 99.83685446303647 * 0.0625
TINT = 128


def _vertex(i, j, freq):
    """World (x, z) of skewed-lattice vertex (i, j)."""
    t = (i + j) * G2
    return (i - t) / freq, (j - t) / freq


def _lattice():
    freq, seed = TINT_TEMP_FREQ, 1337 + 7
    rng = np.random.default_rng(2)
    worst_zero = 0.0
    for _ in range(60):
        i, j = (int(v) for v in rng.integers(-40, 40, 2))
        x, z = _vertex(i, j, freq)
        worst_zero = max(worst_zero, abs(O.biome_noise(seed, freq, x, z)))
        h = 0.01 / freq  # 0.01 noise units
        n0 = O.biome_noise(seed, freq, x, z)
        gx = (O.biome_noise(seed, freq, x + h, z) - O.biome_noise(seed, freq, x - h, z)) / (2 * 0.01)
        gz = (O.biome_noise(seed, freq, x, z + h) - O.biome_noise(seed, freq, x, z - h)) / (2 * 0.01)
        mag = math.hypot(gx, gz)
        assert abs(mag - KERNEL) < 0.02 * KERNEL, (i, j, mag)
        ang = math.degrees(math.atan2(gz, gx)) % 15.0
        assert abs(ang - 7.5) < 0.5, (i, j, ang)
        for k in range(8):
            a = math.radians(45.0 * k + 10.0)
            u = (math.cos(a), math.sin(a))
            d = (O.biome_noise(seed, freq, x + h * u[0], z + h * u[1]) - n0) / 0.01
            want = gx * u[0] + gz * u[1]
            assert abs(d - want) < 0.03 * KERNEL, (i, j, k, d, want)
    assert worst_zero < 2e-4, worst_zero


def test_noise_vanishes_at_vertices_with_a_table_gradient():
    _lattice()


def test_noise_extrema_approach_the_bound():
    xs = np.arange(0, 40000, 53.0)
    zs = np.arange(0, 40000, 61.0)
    v = np.array([O.biome_noise(15, TINT_HUMID_FREQ, x, z) for x in xs[::3] for z in zs[::3]])
    assert np.abs(v).max() <= 1.0005 and np.abs(v).max() > 0.97 and abs(v.mean()) < 0.05


def _smooth(e0, e1, x):
    t = min(max((x - e0) / (e1 - e0), 0.0), 1.0)
    return t * t * (3.0 - 2.0 * t)


TABLE = {"plains": (1.0, 1.0, 1.0), "balanced": (0.935, 1.067, 0.892), "forest": (0.664, 1.045, 1.0),
         "savanna": (1.842, 0.935, 0.793), "desert": (2.096, 0.957, 0.699), "taiga": (0.832, 0.935, 2.096),
         "tundra": (0.755, 0.893, 2.812)}


def _weights(n):
    hot = _smooth(0.04, 0.40, n)
    cold = 1.0 - _smooth(-0.40, -0.04, n)
    return np.array([cold, 1.0 - cold - hot, hot])


def _tint_table():
    from voxelcraft_menu_amd import _abi as A
    rng = np.random.default_rng(5)
    worst, hot_cold = 0.0, set()
    for balance in (0, 1):
        fr = A.Frame()
        set_tint(fr, 1337)
        fb = bytes(fr)
        plains = np.array(TABLE["balanced" if balance else "plains"])
        for _ in range(300):
            x, z = (float(v) for v in rng.uniform(-30000, 30000, 2))
            t = O.biome_noise(1337 + 7, TINT_TEMP_FREQ, x, z)
            h = O.biome_noise(1337 + 8, TINT_HUMID_FREQ, x, z)
            wt, wh = _weights(t), _weights(h)
            T = {k: np.array(v) for k, v in TABLE.items()}
            dry = wt[0] * T["tundra"] + wt[1] * plains + wt[2] * T["desert"]
            mid = wt[0] * T["taiga"] + wt[1] * plains + wt[2] * T["savanna"]
            wet = wt[0] * T["taiga"] + wt[1] * T["forest"] + wt[2] * T["forest"]
            want = wh[0] * dry + wh[1] * mid + wh[2] * wet
            got = O.biome_tint(fb, x, z, balance)
            worst = max(worst, float(np.abs(got - want).max()))
            hot_cold.add((int(np.argmax(wt)), int(np.argmax(wh))))
    assert worst < 2e-5, worst
    assert len(hot_cold) >= 6  # most (temperature, humidity) band pairs were sampled


def test_tint_is_the_table_mix_of_the_band_weights():
    _tint_table()


def _atlas_alpha(alpha):
    a = RP._atlas().reshape(-1, 4).copy()
    a[:, 3] = alpha
    return a.reshape(-1)


def _pillar_tint(strength, alpha=128, legacy=False):
    dense, v, sc, vis, sun, flags, spec_hi = RP._scene(legacy)
    v.frame.flags |= TINT
    set_tint(v.frame, 1337, strength)
    out = O.resolve(sc, v.frame_bytes, O.rspec_from(flags | TINT, spec_hi), _atlas_alpha(alpha), RP.SIZE, RP.LAYERS,
                    RP.MIPS, RP._faces(), vis, sun)
    return dense, v, vis, sun, out


def test_resolve_tint_strength_zero_changes_nothing():
    dense, v, sc, vis, sun, flags, spec_hi = RP._scene()
    base = O.resolve(sc, v.frame_bytes, O.rspec_from(flags, spec_hi), _atlas_alpha(128), RP.SIZE, RP.LAYERS, RP.MIPS,
                     RP._faces(), vis, sun)
    *_, zero = _pillar_tint(0.0)
    assert np.array_equal(zero.view(np.uint32), base.view(np.uint32))


def _floor_tint_misses(legacy=False):
    strength, alpha = 0.8, 128
    dense, v, vis, sun, out = _pillar_tint(strength, alpha, legacy)
    out = out.reshape(-1, 4)
    fb = v.frame_bytes
    lin = [RP._srgb64(c) for c in RP.LEVEL_RGB]
    k = alpha / 255.0 * strength
    bad = n = 0
    for i, t, hit, vox, rd in RP._floor_pixels(dense, v, vis):
        if not RP._open(dense, vox[0], vox[2]):
            continue
        lit = 0.0 if float(np.array([sun[i]], np.uint16).view(np.float16)[0]) > 0 else 1.0
        lod = RP._lod(v, t, rd[1])
        l0 = int(math.floor(lod))
        l1 = min(l0 + 1, RP.MIPS - 1)
        albedo = lin[l0] * (1.0 - (lod - l0)) + lin[l1] * (lod - l0)
        tint = O.biome_tint(fb, float(np.float32(hit[0])), float(np.float32(hit[2]))).astype(np.float64)
        want = albedo * (1.0 - k + tint * k) * RP._light(v, lit, 1.0, legacy)
        n += 1
        if np.abs(out[i, :3] - want).max() > 2e-5 * max(1.0, np.abs(want).max()):
            bad += 1
    assert n > 400, n
    return bad


@pytest.mark.parametrize("legacy", [False, True])
def test_floor_pixels_with_the_tint_closed_form(legacy):
    assert _floor_tint_misses(legacy) == 0


def _fails(fn, *args):
    try:
        fn(*args)
    except AssertionError:
        return True
    return False


@pytest.mark.parametrize("plant,check", [
    (53, "lattice"),   # the skew and unskew factors swapped
    (54, "lattice"),   # the middle corner's kernel dropped
    (55, "table"),     # the temperature bands reversed
    (56, "floor"),     # the tint applied without the texel alpha
])
def test_planted_tint_mistakes_are_caught(plant, check, tmp_path):
    with O.planted(plant, tmp_path):
        if check == "lattice":
            assert _fails(_lattice)
        elif check == "table":
            assert _fails(_tint_table)
        else:
            assert _fails(_floor_tint_misses) or _floor_tint_misses() > 0
