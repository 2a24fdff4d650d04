"""Seeded synthetic clips for the VNLB workloads.

Luma: three sinusoidal gratings of random orientation and period 4-32 px,
scaled into [0, 1], with six random constant-value rectangles painted on
top.  Motion: a global integer translation of 1 px per frame, or a random
per-frame sub-pixel translation of up to 2 px resampled bilinearly.
Chroma (YUV444): two independent low-frequency gratings of amplitude 0.1
around 0.5.  Noise: i.i.d. Gaussian, sigma given in 8-bit units, not
clipped.  Returns (noisy, clean) float32 arrays shaped (F, C, H, W).
"""
from __future__ import annotations

import numpy as np


def _luma_canvas(rng: np.random.Generator, width: int, height: int) -> np.ndarray:
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    canvas = np.zeros((height, width))
    for _ in range(3):
        theta = rng.uniform(0.0, np.pi)
        period = rng.uniform(4.0, 32.0)
        phase = rng.uniform(0.0, 2.0 * np.pi)
        canvas += np.sin(2.0 * np.pi * (xx * np.cos(theta) + yy * np.sin(theta)) / period + phase)
    canvas = (canvas / 3.0 + 1.0) * 0.5
    for _ in range(6):
        rw = int(rng.integers(max(2, width // 16), max(3, width // 3)))
        rh = int(rng.integers(max(2, height // 16), max(3, height // 3)))
        x0 = int(rng.integers(0, max(1, width - rw)))
        y0 = int(rng.integers(0, max(1, height - rh)))
        canvas[y0:y0 + rh, x0:x0 + rw] = rng.uniform(0.0, 1.0)
    return canvas


def _chroma_canvas(rng: np.random.Generator, width: int, height: int) -> np.ndarray:
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    theta = rng.uniform(0.0, np.pi)
    period = rng.uniform(64.0, 256.0)
    phase = rng.uniform(0.0, 2.0 * np.pi)
    return 0.5 + 0.1 * np.sin(2.0 * np.pi * (xx * np.cos(theta) + yy * np.sin(theta)) / period + phase)


def _sample(canvas: np.ndarray, ox: float, oy: float, width: int, height: int, margin: int) -> np.ndarray:
    """Bilinear sample of canvas at (x + margin + ox, y + margin + oy)."""
    fx = np.arange(width) + margin + ox
    fy = np.arange(height) + margin + oy
    x0 = np.floor(fx).astype(np.int64)
    y0 = np.floor(fy).astype(np.int64)
    ax = (fx - x0)[None, :]
    ay = (fy - y0)[:, None]
    c00 = canvas[np.ix_(y0, x0)]
    c01 = canvas[np.ix_(y0, x0 + 1)]
    c10 = canvas[np.ix_(y0 + 1, x0)]
    c11 = canvas[np.ix_(y0 + 1, x0 + 1)]
    return (c00 * (1 - ax) + c01 * ax) * (1 - ay) + (c10 * (1 - ax) + c11 * ax) * ay


def make_clip(width: int, height: int, frames: int, channels: int = 1, sigma8: float = 10.0,
              seed: int = 0, motion: str = "integer") -> tuple[np.ndarray, np.ndarray]:
    if channels not in (1, 3):
        raise ValueError("channels must be 1 or 3")
    rng = np.random.default_rng(seed)
    if motion == "integer":
        offsets = [(-float(t), 0.0) for t in range(frames)]
    elif motion == "subpixel":
        offsets = [tuple(rng.uniform(-2.0, 2.0, size=2)) for _ in range(frames)]
    else:
        raise ValueError("motion must be 'integer' or 'subpixel'")
    margin = int(np.ceil(max(max(abs(o[0]), abs(o[1])) for o in offsets))) + 2
    cw, ch = width + 2 * margin + 2, height + 2 * margin + 2
    canvases = [_luma_canvas(rng, cw, ch)]
    if channels == 3:
        canvases += [_chroma_canvas(rng, cw, ch), _chroma_canvas(rng, cw, ch)]
    clean = np.empty((frames, channels, height, width), dtype=np.float32)
    for t, (ox, oy) in enumerate(offsets):
        for c, canvas in enumerate(canvases):
            clean[t, c] = _sample(canvas, ox, oy, width, height, margin)
    noise = rng.standard_normal(clean.shape) * (sigma8 / 255.0)
    noisy = (clean + noise).astype(np.float32)
    return noisy, clean


def psnr(a: np.ndarray, b: np.ndarray, peak: float = 1.0) -> float:
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0.0 else 10.0 * np.log10(peak * peak / mse)
