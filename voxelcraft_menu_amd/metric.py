"""Image diff metrics of the reference's harness (src/harness/metric.rs:102-240) over RGBA8
images as (H, W, 4) uint8 arrays (what Renderer.read_image returns), so presented frames
can be compared with the reference's own measures.

A rect is (x0, y0, x1, y1), clamped to the image as Rect::clamped does (metric.rs:75-82);
None is the whole image. Only the first three channels are compared.
"""
from __future__ import annotations

import numpy as np


def _rect(img: np.ndarray, r) -> tuple[int, int, int, int]:
    h, w = img.shape[:2]
    if r is None:
        return 0, 0, w, h
    x0, y0, x1, y1 = r
    return min(x0, w), min(y0, h), min(x1, w), min(y1, h)


def _pixels(r) -> int:
    x0, y0, x1, y1 = r
    return max(x1 - x0, 0) * max(y1 - y0, 0)


def _crop(img: np.ndarray, r) -> np.ndarray:
    x0, y0, x1, y1 = r
    return img[y0:y1, x0:x1, :3].astype(np.int64)


def mae(a: np.ndarray, b: np.ndarray, rect=None) -> float:
    """Mean absolute difference per channel (metric.rs:102-114)."""
    r = _rect(a, rect)
    s = np.abs(_crop(a, r) - _crop(b, r)).sum()
    return float(s) / (max(_pixels(r), 1) * 3.0)


def max_delta(a: np.ndarray, b: np.ndarray, rect=None) -> int:
    """Largest channel difference (metric.rs:116-128)."""
    r = _rect(a, rect)
    d = np.abs(_crop(a, r) - _crop(b, r))
    return int(d.max()) if d.size else 0


def diff_bbox(a: np.ndarray, b: np.ndarray):
    """Bounding rect of the differing pixels, or None (metric.rs:130-152)."""
    if a.shape[:2] != b.shape[:2]:
        return None
    diff = (a[..., :3] != b[..., :3]).any(axis=-1)
    ys, xs = np.nonzero(diff)
    if xs.size == 0:
        return None
    return int(xs.min()), int(ys.min()), int(xs.max()) + 1, int(ys.max()) + 1


def pixels_differing(a: np.ndarray, b: np.ndarray, rect=None) -> int:
    """Pixels with any channel different (metric.rs:154-165)."""
    r = _rect(a, rect)
    return int((_crop(a, r) != _crop(b, r)).any(axis=-1).sum())


def speckle(a: np.ndarray, rect=None) -> float:
    """Mean absolute deviation of each channel from its 3x3 box mean, edges clamped
    (metric.rs:167-186)."""
    r = _rect(a, rect)
    x0, y0, x1, y1 = r
    img = a[..., :3].astype(np.float64)
    pad = np.pad(img, ((1, 1), (1, 1), (0, 0)), mode="edge")
    h, w = img.shape[:2]
    m = sum(pad[1 + dy:1 + dy + h, 1 + dx:1 + dx + w] for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    dev = np.abs(img - m / 9.0)[y0:y1, x0:x1]
    return float(dev.sum()) / (max(_pixels(r), 1) * 3.0)


def mean_rgb(a: np.ndarray, rect=None) -> list[float]:
    """Mean of each colour channel (metric.rs:226-240)."""
    r = _rect(a, rect)
    s = _crop(a, r).reshape(-1, 3).sum(axis=0)
    n = max(_pixels(r), 1)
    return [float(s[c]) / n for c in range(3)]
