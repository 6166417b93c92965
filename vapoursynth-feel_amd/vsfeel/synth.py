"""Seeded synthetic clips for the BM3Dv2 workloads.

Every frame is a clean image in [0, 1] built from a linear gradient, 20
axis-aligned rectangles of uniform intensity and a sinusoidal texture
(amplitude 0.1, period 6-20 px), plus i.i.d. Gaussian noise, clamped to
[0, 1] and stored as float32.  The choices the workload description leaves
open are frozen here:

* gradient: ``a + bx * x / W + by * y / H`` with ``a ~ U(0.15, 0.35)`` and
  ``bx, by ~ U(-0.25, 0.35)``;
* rectangles: corner ``~ U`` over the frame, sides ``~ U(W/40, W/6)`` by
  ``U(H/40, H/6)`` (at least 2 px), painted in order with an intensity
  ``~ U(0.05, 0.95)``;
* texture: ``0.1 * sin(2 pi (x cos t + y sin t) / period + phase)`` with
  ``period ~ U(6, 20)``, ``t ~ U(0, pi)``, ``phase ~ U(0, 2 pi)``;
* seeds: numpy ``PCG64(1000 * config + frame)``; a clip whose frames are
  shifted copies of one image draws the image from ``PCG64(1000 * config)``
  and the noise of frame t from ``PCG64(1000 * config + 1 + t)``.
"""

from __future__ import annotations

import numpy as np

__all__ = [
    "clean_image",
    "noisy_frame",
    "gray_clip",
    "shifted_clip",
    "smooth_gradient",
    "workload",
]


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def clean_image(rng: np.random.Generator, width: int, height: int) -> np.ndarray:
    """One clean float64 image in [0, 1] (before noise and clamping)."""
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    a = rng.uniform(0.15, 0.35)
    bx, by = rng.uniform(-0.25, 0.35, size=2)
    img = a + bx * x / width + by * y / height
    for _ in range(20):
        w = max(2, int(rng.uniform(width / 40, width / 6)))
        h = max(2, int(rng.uniform(height / 40, height / 6)))
        x0 = int(rng.integers(0, max(1, width - 1)))
        y0 = int(rng.integers(0, max(1, height - 1)))
        img[y0:y0 + h, x0:x0 + w] = rng.uniform(0.05, 0.95)
    period = rng.uniform(6.0, 20.0)
    theta = rng.uniform(0.0, np.pi)
    phase = rng.uniform(0.0, 2.0 * np.pi)
    img += 0.1 * np.sin(2.0 * np.pi * (x * np.cos(theta) + y * np.sin(theta)) / period + phase)
    return img


def _finish(img: np.ndarray, rng: np.random.Generator, noise_std: float) -> np.ndarray:
    if noise_std > 0:
        img = img + rng.normal(0.0, noise_std, size=img.shape)
    return np.clip(img, 0.0, 1.0).astype(np.float32)


def noisy_frame(config: int, frame: int, width: int, height: int, noise_std: float) -> np.ndarray:
    rng = _rng(1000 * config + frame)
    return _finish(clean_image(rng, width, height), rng, noise_std)


def gray_clip(config: int, nframes: int, width: int, height: int, noise_std: float) -> np.ndarray:
    """(nframes, height, width) float32: an independent image per frame."""
    out = np.empty((nframes, height, width), np.float32)
    for t in range(nframes):
        out[t] = noisy_frame(config, t, width, height, noise_std)
    return out


def _bilinear_shift(img: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """img sampled at (x - dx, y - dy), edges clamped."""
    h, w = img.shape
    ys = np.arange(h, dtype=np.float64)[:, None] - dy
    xs = np.arange(w, dtype=np.float64)[None, :] - dx
    ys = np.clip(ys, 0.0, h - 1.0)
    xs = np.clip(xs, 0.0, w - 1.0)
    y0 = np.floor(ys).astype(np.int64)
    x0 = np.floor(xs).astype(np.int64)
    y1 = np.minimum(y0 + 1, h - 1)
    x1 = np.minimum(x0 + 1, w - 1)
    fy = ys - y0
    fx = xs - x0
    top = img[y0, x0] * (1.0 - fx) + img[y0, x1] * fx
    bot = img[y1, x0] * (1.0 - fx) + img[y1, x1] * fx
    return top * (1.0 - fy) + bot * fy


def shifted_clip(config: int, nframes: int, width: int, height: int, noise_std: float,
                 shift: tuple[float, float]) -> np.ndarray:
    """Frame t is the config's image moved by (shift[0] * t, shift[1] * t) px."""
    base = clean_image(_rng(1000 * config), width, height)
    out = np.empty((nframes, height, width), np.float32)
    for t in range(nframes):
        moved = _bilinear_shift(base, shift[0] * t, shift[1] * t)
        out[t] = _finish(moved, _rng(1000 * config + 1 + t), noise_std)
    return out


def smooth_gradient(seed: int, width: int, height: int, noise_std: float) -> np.ndarray:
    """A chroma-like plane: a smooth 2-D gradient plus noise, float32."""
    rng = _rng(seed)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    c = rng.uniform(0.3, 0.7, size=6)
    img = c[0] + (c[1] - 0.5) * x / width + (c[2] - 0.5) * y / height
    img += 0.1 * np.sin(np.pi * (c[3] * x / width + c[4] * y / height) + 6.0 * c[5])
    return _finish(img, rng, noise_std)


# Workload shapes and filter arguments, by configuration number.
WORKLOADS = {
    1: dict(frames=8, width=640, height=360, noise=3 / 255, fmt="gray",
            args=dict(sigma=3.0, block_step=8, bm_range=9, radius=0, ps_num=2, ps_range=4,
                      extractor_exp=0)),
    2: dict(frames=32, width=1920, height=1080, noise=3 / 255, fmt="gray",
            args=dict(sigma=3.0, block_step=4, bm_range=16, radius=0, extractor_exp=0)),
    3: dict(frames=32, width=1920, height=1080, noise=3 / 255, fmt="yuv420",
            shift=(1.5, 0.75),
            args=dict(sigma=3.0, block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=7,
                      extractor_exp=2)),
    4: dict(frames=16, width=3840, height=2160, noise=5 / 255, fmt="gray", shift=(1.0, 0.0),
            args=dict(sigma=5.0, block_step=4, bm_range=16, radius=1, ps_num=2, ps_range=7)),
    5: dict(frames=16, width=512, height=512, noise=3 / 255, fmt="gray",
            args=dict(sigma=3.0, block_step=4, bm_range=9, radius=0)),
}


def workload(config: int, frames: int | None = None, width: int | None = None,
             height: int | None = None):
    """Planes of a workload: a (N,H,W) luma array, plus (U, V) for YUV420.

    frames / width / height override the workload's own sizes (the same
    generator at a smaller size, for tests the oracle can finish quickly).
    """
    spec = WORKLOADS[config]
    n = spec["frames"] if frames is None else frames
    w = spec["width"] if width is None else width
    h = spec["height"] if height is None else height
    if "shift" in spec:
        luma = shifted_clip(config, n, w, h, spec["noise"], spec["shift"])
    else:
        luma = gray_clip(config, n, w, h, spec["noise"])
    if spec["fmt"] != "yuv420":
        return luma
    cw, ch = w // 2, h // 2
    u = np.stack([smooth_gradient(1000 * config + 500 + t, cw, ch, spec["noise"]) for t in range(n)])
    v = np.stack([smooth_gradient(1000 * config + 700 + t, cw, ch, spec["noise"]) for t in range(n)])
    return luma, u, v
