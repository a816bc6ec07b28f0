"""The oracle's present blit (blit.wgsl:18-93) against float64 closed forms, and the image
metrics (src/harness/metric.rs) on hand-made images.

Present: below the knee (0.75) the tonemap is the identity and above it
1 - 0.0625 / (x - 0.5); tone 1 is the ACES fit clamp(x(2.51x + 0.03) / (x(2.43x + 0.59)
+ 0.14)); grade 1 moves each channel away from the luma (0.2126, 0.7152, 0.0722) by 1.06
after a (1.05, 1, 0.93) balance; the sRGB encode is 12.92 c below 0.0031308 and
1.055 c^(1/2.4) - 0.055 above; the byte is round(255 c). The float64 value and the byte
may differ by one only where the float64 value lies within 1e-4 of a rounding boundary.
"""
import numpy as np
import pytest

from oracle import oracle as O
from voxelcraft_menu_amd import metric as M


def _srgb(c):
    return np.where(c <= 0.0031308, c * 12.92, 1.055 * np.maximum(c, 0.0031308) ** (1 / 2.4) - 0.055)


def _tone(x):
    return np.where(x < 0.75, x, 1.0 - 0.0625 / (np.maximum(x, 0.75) - 0.5))


def _aces(x):
    return np.clip(x * (2.51 * x + 0.03) / (x * (2.43 * x + 0.59) + 0.14), 0.0, 1.0)


def _grade1(c):
    luma = (c * np.array([0.2126, 0.7152, 0.0722])).sum(-1, keepdims=True)
    bal = np.clip(c * np.array([1.05, 1.0, 0.93]), 0, 1)
    return np.clip(luma + (bal - luma) * 1.06, 0, 1)


def _bytes_match(got, want):
    w255 = np.clip(want, 0, 1) * 255.0
    exact = np.floor(w255 + 0.5)
    near = np.abs(w255 - np.floor(w255) - 0.5) < 1e-4 * 255
    ok = (got == exact) | (near & (np.abs(got - exact) <= 1))
    return bool(ok.all()), int((~ok).sum())


def _image(seed=1, n=4096):
    rng = np.random.default_rng(seed)
    img = np.zeros((n // 64, 64, 4), np.float32)
    img[..., :3] = np.concatenate([rng.uniform(0, 1.5, (n // 2, 3)), rng.uniform(-0.2, 8.0, (n // 2, 3))]) \
        .reshape(n // 64, 64, 3)
    img[..., 3] = 1.0
    return img


@pytest.mark.parametrize("encode,tone,grade", [(1, 0, 0), (0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)])
def test_present_closed_forms(encode, tone, grade):
    src = _image()
    got = O.present(src, encode, tone, grade, 1.0)
    hdr = np.maximum(src[..., :3].astype(np.float64), 0.0)
    c = _aces(hdr) if tone == 1 else _tone(hdr)
    if grade == 1:
        c = _grade1(c)
    if encode:
        c = _srgb(c)
    ok, bad = _bytes_match(got[..., :3].astype(np.float64), c)
    assert ok, bad
    assert (got[..., 3] == 255).all()


def test_present_grade_strength_mixes():
    src = _image(2)
    full = O.present(src, 0, 0, 1, 1.0).astype(np.float64)
    none = O.present(src, 0, 0, 1, 0.0).astype(np.float64)
    plain = O.present(src, 0, 0, 0, 1.0).astype(np.float64)
    assert np.array_equal(none, plain)  # strength 0: the ungraded colour
    half = O.present(src, 0, 0, 1, 0.5).astype(np.float64)
    assert np.abs(half - (full + plain) / 2).max() <= 1.0


def test_present_grade_two_is_finite_and_bounded():
    src = _image(3)
    out = O.present(src, 1, 0, 2, 0.7)
    assert out.dtype == np.uint8 and (out[..., 3] == 255).all()
    g = O.present(src, 0, 0, 2, 1.0).astype(np.int64) - O.present(src, 0, 0, 0, 1.0).astype(np.int64)
    assert np.abs(g[..., :3]).max() > 0  # the grade changes the image


def test_metrics_on_hand_made_images():
    a = np.zeros((10, 20, 4), np.uint8)
    b = a.copy()
    b[..., :3] = 3
    assert M.mae(a, b) == 3.0 and M.max_delta(a, b) == 3
    assert M.pixels_differing(a, b) == 200
    c = a.copy()
    c[4, 7, 1] = 200
    c[6, 9, 0] = 10
    assert M.diff_bbox(a, c) == (7, 4, 10, 7)
    assert M.diff_bbox(a, a) is None
    assert M.pixels_differing(a, c) == 2 and M.max_delta(a, c) == 200
    assert M.mae(a, c) == pytest.approx(210 / (200 * 3))
    assert M.mae(a, c, (7, 4, 8, 5)) == pytest.approx(200 / 3)
    assert M.pixels_differing(a, c, (0, 0, 100, 5)) == 1  # the rect is clamped to the image
    assert M.speckle(a) == 0.0
    d = a.copy()
    d[5, 10, 0] = 90  # a single bright sample: |90 - 10| at itself, |0 - 10| at its 8 neighbours
    assert M.speckle(d) == pytest.approx((80 + 8 * 10) / (200 * 3))
    assert M.mean_rgb(b) == [3.0, 3.0, 3.0]
