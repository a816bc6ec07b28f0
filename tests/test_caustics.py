"""Caustics (resolve.wgsl:530-549, SPEC_CAUSTICS) in the oracle, pinned by a float64 closed
form: under a sea level 4 blocks above the pillar chunk's floor, each open floor pixel's
sun term is multiplied by 1 + 4000 r0 r1 exp(-0.3 d), r = |g|^2 of the first two wave
octaves (wave_octave without anisotropy: g = dir slope fade cos(2 pi / lambda <p, dir> +
phase), fade = smoothstep(2, 4, lambda / 0.35)) at the surface point above the hit along
the sun, d the depth. A planted mistake (the sheets summed instead of multiplied) fails it.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import test_oracle_resolve_pins as RP

CAUSTICS = 64


def _smooth(e0, e1, x):
    t = min(max((x - e0) / (e1 - e0), 0.0), 1.0)
    return t * t * (3.0 - 2.0 * t)


def _octave(p, d, lam, slope, phase):
    fade = _smooth(2.0, 4.0, lam / 0.35)
    g = slope * fade * math.cos(2.0 * math.pi / lam * (p[0] * d[0] + p[1] * d[1]) + phase)
    return (d[0] * g) ** 2 + (d[1] * g) ** 2


def _misses():
    dense, v, sc, vis, sun, flags, spec_hi = RP._scene()
    f = v.frame
    f.sea_level = 24.0
    f.wave_amp, f.wave_scale, f.wave_speed, f.time = 0.5, 16.0, 0.3, 1.7
    out = O.resolve(sc, v.frame_bytes, O.rspec_from(flags, spec_hi | CAUSTICS), RP._atlas(), RP.SIZE, RP.LAYERS,
                    RP.MIPS, RP._faces(), vis, sun).reshape(-1, 4)
    d, sy = float(f.daylight), float(np.float32(f.sun_dir[1]))
    sx, sz = float(np.float32(f.sun_dir[0])), float(np.float32(f.sun_dir[2]))
    lin = [RP._srgb64(c) for c in RP.LEVEL_RGB]
    a, l0 = 0.5, 16.0
    ph = 1.7 * 0.3 * 6.28318531
    bad = n = lifted = 0
    for i, t, hit, vox, rd in RP._floor_pixels(dense, v, vis):
        if not RP._open(dense, vox[0], vox[2]):
            continue
        lit = 0.0 if float(np.array([sun[i]], np.uint16).view(np.float16)[0]) > 0 else 1.0
        depth = 24.0 - 20.0
        run = depth / max(sy, 1e-3)
        p = (hit[0] - sx * run, hit[2] - sz * run)
        r0 = _octave(p, (0.913089, 0.407760), l0, a * 0.42, ph)
        r1 = _octave(p, (-0.388685, 0.921371), l0 / 2.17, a * 0.27, ph * 1.4731)
        gain = 1.0 + r0 * r1 * 4000.0 * math.exp(-depth * 0.30)
        direct = lit * sy * d * gain if sy > 0.0 and d > 0.01 else 0.0
        light = 0.5 * RP.SKY_TINT * 0.6 * d + RP.SUN_RGB * direct * 0.78 + RP.FLOOR_RGB * float(f.ambient)
        lod = RP._lod(v, t, rd[1])
        l_0 = int(math.floor(lod))
        l_1 = min(l_0 + 1, RP.MIPS - 1)
        albedo = lin[l_0] * (1.0 - (lod - l_0)) + lin[l_1] * (lod - l_0)
        want = albedo * light
        n += 1
        lifted += lit > 0 and gain > 1.05
        if np.abs(out[i, :3] - want).max() > 5e-5 * max(1.0, np.abs(want).max()):
            bad += 1
    assert n > 400 and lifted > 100, (n, lifted)
    return bad


def test_floor_caustics_closed_form():
    assert _misses() == 0


def test_planted_caustic_mistake_is_caught(tmp_path):
    with O.planted(62, tmp_path):
        try:
            bad = _misses()
        except AssertionError:
            bad = 1
        assert bad > 0
