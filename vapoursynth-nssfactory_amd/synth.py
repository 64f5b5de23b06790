"""Seeded synthetic inputs for the BM3D workloads (numpy PCG64).

Frozen generator specification (the parameters the workload text leaves open
are chosen here and recorded):

luma / gray clean image (``clean_image``):
    0.5 + sum over 6 sinusoids A_i * sin(2 pi (fx_i x + fy_i y) + phi_i),
    A_i ~ U[0.04, 0.12], |f_i| ~ U[0.01, 0.2] cycles/px at a uniform angle,
    phi_i ~ U[0, 2 pi); plus 20 axis-aligned constant rectangles, side
    lengths ~ U[min(W,H)/32, min(W,H)/4] px, value ~ U[0, 1], drawn in
    order (later ones on top); clipped to [0, 1].
chroma plane (``chroma_field``): 0.5 * (sum of 3 sinusoids with
    |f| ~ U[0.002, 0.02] cycles/px, unit amplitude, divided by 3), in
    [-0.5, 0.5].
noise: i.i.d. N(0, (sigma/255)^2) per plane, float32.
"""
from __future__ import annotations

import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def clean_image(width: int, height: int, seed: int = 0) -> np.ndarray:
    rng = _rng(seed)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    img = np.full((height, width), 0.5)
    for _ in range(6):
        a = rng.uniform(0.04, 0.12)
        f = rng.uniform(0.01, 0.2)
        ang = rng.uniform(0, 2 * np.pi)
        phi = rng.uniform(0, 2 * np.pi)
        img += a * np.sin(2 * np.pi * (f * np.cos(ang) * x + f * np.sin(ang) * y) + phi)
    m = min(width, height)
    lo, hi = max(1, m // 32), max(2, m // 4)
    for _ in range(20):
        rw = int(rng.integers(lo, hi + 1))
        rh = int(rng.integers(lo, hi + 1))
        x0 = int(rng.integers(0, max(1, width - rw + 1)))
        y0 = int(rng.integers(0, max(1, height - rh + 1)))
        img[y0:y0 + rh, x0:x0 + rw] = rng.uniform(0, 1)
    return np.clip(img, 0, 1).astype(np.float32)


def chroma_field(width: int, height: int, seed: int = 0) -> np.ndarray:
    rng = _rng(seed)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    acc = np.zeros((height, width))
    for _ in range(3):
        f = rng.uniform(0.002, 0.02)
        ang = rng.uniform(0, 2 * np.pi)
        phi = rng.uniform(0, 2 * np.pi)
        acc += np.sin(2 * np.pi * (f * np.cos(ang) * x + f * np.sin(ang) * y) + phi)
    return (0.5 * acc / 3.0).astype(np.float32)


def add_noise(plane: np.ndarray, sigma_user: float, seed: int) -> np.ndarray:
    rng = _rng(seed)
    n = rng.standard_normal(plane.shape) * (sigma_user / 255.0)
    return (plane.astype(np.float64) + n).astype(np.float32)


FORMATS = {
    # name: (num_planes, subsampling_w, subsampling_h, color family)
    "GRAYS": (1, 0, 0, "gray"),
    "YUV420PS": (3, 1, 1, "yuv"),
    "YUV444PS": (3, 0, 0, "yuv"),
    "RGBS": (3, 0, 0, "rgb"),
}


def plane_dims(fmt: str, width: int, height: int):
    n, sw, sh, _ = FORMATS[fmt]
    return [(width, height) if i == 0 else (width >> sw, height >> sh) for i in range(n)]


def clean_frame(fmt: str, width: int, height: int, seed: int = 0):
    """Clean planes of one frame: luma/gray from ``clean_image``; YUV chroma
    from ``chroma_field``; RGB planes are three luma-style images."""
    _, _, _, family = FORMATS[fmt]
    dims = plane_dims(fmt, width, height)
    out = []
    for i, (w, h) in enumerate(dims):
        if i == 0 or family == "rgb":
            out.append(clean_image(w, h, seed * 16 + i))
        else:
            out.append(chroma_field(w, h, seed * 16 + i))
    return out


def noisy_frame(fmt: str, width: int, height: int, sigma_user: float, seed: int = 0):
    clean = clean_frame(fmt, width, height, seed)
    return clean, [add_noise(p, sigma_user, 1000 + seed * 16 + i) for i, p in enumerate(clean)]


def shifted_sequence(fmt: str, width: int, height: int, frames: int, sigma_user: float, seed: int = 0,
                     max_step: int = 4):
    """Base frame shifted by a cumulative random integer translation per
    frame (step ~ U{-max_step..max_step} in x and y, edge-replicated; chroma
    planes shift by the luma shift scaled to their subsampling), with fresh
    noise per frame."""
    rng = _rng(seed + 7)
    base = clean_frame(fmt, width, height, seed)
    n, sw, sh, _ = FORMATS[fmt]
    dx = dy = 0
    clean_seq, noisy_seq = [], []
    for f in range(frames):
        if f:
            dx += int(rng.integers(-max_step, max_step + 1))
            dy += int(rng.integers(-max_step, max_step + 1))
        planes = []
        for i, p in enumerate(base):
            sx = dx >> sw if i else dx
            sy = dy >> sh if i else dy
            h, w = p.shape
            yi = np.clip(np.arange(h) - sy, 0, h - 1)
            xi = np.clip(np.arange(w) - sx, 0, w - 1)
            planes.append(np.ascontiguousarray(p[yi][:, xi]))
        clean_seq.append(planes)
        noisy_seq.append([add_noise(p, sigma_user, 50000 + f * 16 + i + seed * 7919) for i, p in enumerate(planes)])
    return clean_seq, noisy_seq
