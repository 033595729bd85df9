"""Seeded synthetic speckle scenes for the 2D-DIC workloads.

Every scene is generated from a seed with numpy's PCG64, so the same call
returns the same bytes on any machine with the same numpy.  A scene holds the
reference and target images, the POI queue and a truth function giving the
exact displacement and gradients at a reference point.

Speckle: Gaussian dots with centres uniform over the image, radius
R ~ U[2, 4] px rendered with sigma = R / 2; intensity
background + amplitude * min(1, sum of dot profiles), so overlapping dots
saturate at background + amplitude; additive Gaussian noise; rounded and
clipped to the pixel range.

Target: the noisy reference resampled with Keys cubic convolution (a = -0.5,
edge-clamped) at the back-projected position X(x) of every target pixel, where
x = X + u(X) is the forward displacement map, then rounded and clipped.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Callable

import numpy as np

from ._native import POI_DTYPE
from .opencorr import grid_queue


@dataclass
class Scene:
    name: str
    ref: np.ndarray
    tar: np.ndarray
    pois: np.ndarray
    radius: int
    truth: Callable[[np.ndarray, np.ndarray], tuple]  # (X, Y) -> (u, v, ux, uy, vx, vy)
    meta: dict = field(default_factory=dict)
    depth12: str = "native"


def dot_field(width: int, height: int, n_dots: int, rng: np.random.Generator,
              radius=(2.0, 4.0)):
    cx = rng.uniform(0.0, width, n_dots)
    cy = rng.uniform(0.0, height, n_dots)
    sigma = rng.uniform(radius[0], radius[1], n_dots) / 2.0
    return cx, cy, sigma


def render_dots(cx, cy, sigma, width: int, height: int, background: float,
                amplitude: float) -> np.ndarray:
    """Noise-free speckle image (float64, height x width), pixels sampled at
    integer coordinates."""
    reach = int(math.ceil(4.0 * float(np.max(sigma)))) if len(sigma) else 0
    acc = np.zeros(width * height, dtype=np.float64)
    fx = np.floor(cx).astype(np.int64)
    fy = np.floor(cy).astype(np.int64)
    inv2s2 = 1.0 / (2.0 * sigma * sigma)
    offsets = [(dx, dy) for dy in range(-reach, reach + 2) for dx in range(-reach, reach + 2)]
    chunk = 16
    for k in range(0, len(offsets), chunk):
        idx_parts, w_parts = [], []
        for dx, dy in offsets[k:k + chunk]:
            ix = fx + dx
            iy = fy + dy
            ok = (ix >= 0) & (ix < width) & (iy >= 0) & (iy < height)
            d2 = (ix - cx) ** 2 + (iy - cy) ** 2
            w = np.exp(-d2 * inv2s2)
            ok &= d2 <= (4.0 * sigma) ** 2
            idx_parts.append((iy * width + ix)[ok])
            w_parts.append(w[ok])
        acc += np.bincount(np.concatenate(idx_parts), np.concatenate(w_parts), minlength=acc.size)
    img = background + amplitude * np.minimum(1.0, acc)
    return img.reshape(height, width)


def quantize(img: np.ndarray, rng: np.random.Generator, noise: float, dtype, vmax: float) -> np.ndarray:
    noisy = img + (rng.normal(0.0, noise, img.shape) if noise > 0 else 0.0)
    return np.clip(np.rint(noisy), 0, vmax).astype(dtype)


def _keys(t: np.ndarray):
    a = -0.5
    t1 = t + 1.0
    t2 = 1.0 - t
    t3 = 2.0 - t
    w0 = ((a * t1 - 5 * a) * t1 + 8 * a) * t1 - 4 * a
    w1 = ((a + 2) * t - (a + 3)) * t * t + 1
    w2 = ((a + 2) * t2 - (a + 3)) * t2 * t2 + 1
    w3 = ((a * t3 - 5 * a) * t3 + 8 * a) * t3 - 4 * a
    return w0, w1, w2, w3


def cubic_resample(img: np.ndarray, X: np.ndarray, Y: np.ndarray) -> np.ndarray:
    """Keys cubic convolution of img at (X, Y), edge-clamped."""
    h, w = img.shape
    src = img.astype(np.float64)
    ix = np.floor(X).astype(np.int64)
    iy = np.floor(Y).astype(np.int64)
    wx = _keys(X - ix)
    wy = _keys(Y - iy)
    out = np.zeros(X.shape, dtype=np.float64)
    for j in range(4):
        yy = np.clip(iy + j - 1, 0, h - 1)
        row = np.zeros(X.shape, dtype=np.float64)
        for i in range(4):
            xx = np.clip(ix + i - 1, 0, w - 1)
            row += wx[i] * src[yy, xx]
        out += wy[j] * row
    return out


def warp_target(ref: np.ndarray, inverse_map, dtype, vmax: float, rows_per_chunk: int = 256) -> np.ndarray:
    """Target pixel x takes the reference resampled at X = inverse_map(x)."""
    h, w = ref.shape
    out = np.empty((h, w), dtype=dtype)
    xs = np.arange(w, dtype=np.float64)
    for y0 in range(0, h, rows_per_chunk):
        ys = np.arange(y0, min(h, y0 + rows_per_chunk), dtype=np.float64)
        yy, xx = np.meshgrid(ys, xs, indexing="ij")
        X, Y = inverse_map(xx, yy)
        v = cubic_resample(ref, X, Y)
        out[y0:y0 + len(ys)] = np.clip(np.rint(v), 0, vmax).astype(dtype)
    return out


def affine_scene(name: str, width: int, height: int, n_dots: int, seed: int, t=(0.0, 0.0),
                 grad=(0.0, 0.0, 0.0, 0.0), radius: int = 15, grid: int = 10,
                 background: float = 40.0, peak: float = 180.0, noise: float = 1.5,
                 dtype=np.uint8, vmax: float = 255.0) -> Scene:
    """x = c + A (X - c) + t, A = I + [[ux, uy], [vx, vy]], c the image centre."""
    rng = np.random.default_rng(seed)
    cx, cy, s = dot_field(width, height, n_dots, rng)
    ref = quantize(render_dots(cx, cy, s, width, height, background, peak - background), rng,
                   noise, dtype, vmax)
    ux, uy, vx, vy = grad
    A = np.array([[1.0 + ux, uy], [vx, 1.0 + vy]])
    Ai = np.linalg.inv(A)
    c = ((width - 1) / 2.0, (height - 1) / 2.0)

    def inverse_map(x, y):
        dx, dy = x - c[0] - t[0], y - c[1] - t[1]
        return c[0] + Ai[0, 0] * dx + Ai[0, 1] * dy, c[1] + Ai[1, 0] * dx + Ai[1, 1] * dy

    tar = warp_target(ref, inverse_map, dtype, vmax)

    def truth(X, Y):
        X = np.asarray(X, dtype=np.float64)
        Y = np.asarray(Y, dtype=np.float64)
        u = t[0] + ux * (X - c[0]) + uy * (Y - c[1])
        v = t[1] + vx * (X - c[0]) + vy * (Y - c[1])
        one = np.ones_like(X)
        return u, v, ux * one, uy * one, vx * one, vy * one

    pois = grid_queue(width, height, radius, grid)
    meta = {"seed": seed, "n_dots": n_dots, "translation": list(t), "gradients": list(grad),
            "grid": grid, "margin": radius + 2, "noise_sigma": noise, "background": background,
            "peak": peak, "sigma_rule": "R/2, R~U[2,4]", "resample": "keys-cubic a=-0.5"}
    return Scene(name, ref, tar, pois, radius, truth, meta)


def config0(seed: int = 20260) -> Scene:
    """512x512 u8, ~6000 dots, t = (2.35, -1.70), ux = vy = 1e-3, r = 15, grid 10."""
    return affine_scene("config0", 512, 512, 6000, seed, t=(2.35, -1.70),
                        grad=(1e-3, 0.0, 0.0, 1e-3), radius=15, grid=10)


def config1(seed: int = 20261, size: int = 2048, grid: int = 5, radius: int = 15) -> Scene:
    """size^2 u8 speckle at the dot density of config 0.  Displacement
    u = U + 0.5 sin(2 pi X / 256), v = V + 0.5 cos(2 pi Y / 256), (U, V) drawn
    from U[-20, 20].  Each POI starts from the global integer guess
    (round(U), round(V)) so FFTCC's capture range (-(r-1)..r) holds the
    residual."""
    rng = np.random.default_rng(seed)
    U, V = rng.uniform(-20.0, 20.0, 2)
    n_dots = int(round(6000 * size * size / (512 * 512)))
    cx, cy, s = dot_field(size, size, n_dots, rng)
    ref = quantize(render_dots(cx, cy, s, size, size, 40.0, 140.0), rng, 1.5, np.uint8, 255.0)
    k = 2.0 * math.pi / 256.0

    def fwd(X, Y):
        return U + 0.5 * np.sin(k * X), V + 0.5 * np.cos(k * Y)

    def inverse_map(x, y):
        X, Y = x - U, y - V
        for _ in range(8):
            u, v = fwd(X, Y)
            X, Y = x - u, y - v
        return X, Y

    tar = warp_target(ref, inverse_map, np.uint8, 255.0)

    def truth(X, Y):
        X = np.asarray(X, dtype=np.float64)
        Y = np.asarray(Y, dtype=np.float64)
        u, v = fwd(X, Y)
        z = np.zeros_like(X)
        return u, v, 0.5 * k * np.cos(k * X), z, z, -0.5 * k * np.sin(k * Y)

    pois = grid_queue(size, size, radius, grid)
    pois["deformation"][:, 0] = np.float32(round(U))
    pois["deformation"][:, 6] = np.float32(round(V))
    meta = {"seed": seed, "n_dots": n_dots, "U": float(U), "V": float(V),
            "initial_guess": [round(U), round(V)], "grid": grid, "margin": radius + 2,
            "noise_sigma": 1.5, "sigma_rule": "R/2, R~U[2,4]", "resample": "keys-cubic a=-0.5"}
    return Scene(f"config1_{size}", ref, tar, pois, radius, truth, meta)


RADII_MIXED = (10, 15, 20)


@dataclass
class Sequence:
    """A fixed reference and a family of targets: frame k is the reference
    warped by the affine map of step k (frame(k) renders it)."""
    name: str
    ref: np.ndarray
    pois: np.ndarray
    frames: int
    frame: Callable[[int], np.ndarray]
    truth: Callable[[int, np.ndarray, np.ndarray], tuple]  # (k, X, Y) -> (u, v, ux, uy, vx, vy)
    meta: dict = field(default_factory=dict)
    depth12: str = "right"


def config2(size: int = 1024, frames: int = 500, grid: int = 5, seed: int = 20262,
            radii=RADII_MIXED) -> Sequence:
    """Batched multi-frame sequence: size^2 u16 holding 12-bit values
    (background 640, peak 2880, noise sigma 24 grey levels), one fixed
    reference; target k is the reference resampled (as every target here)
    under the map of frame k (0 <= k < frames) translates by t_k = 15 k / 499 px
    along 30 degrees, (u, v) = t_k (cos 30, sin 30), plus the equibiaxial strain
    eps_k = 5e-3 k / 499 about the image centre (ux = vy = eps_k).  POIs on a
    `grid` px grid at margin max(radii) + 2; POI i has the square subset radius
    radii[i % len(radii)] (self-adaptive subsets)."""
    rng = np.random.default_rng(seed)
    n_dots = int(round(6000 * size * size / (512 * 512)))
    cx, cy, s = dot_field(size, size, n_dots, rng)
    ref = quantize(render_dots(cx, cy, s, size, size, 640.0, 2240.0), rng, 24.0, np.uint16, 4095.0)
    c = ((size - 1) / 2.0, (size - 1) / 2.0)
    dirx, diry = math.cos(math.pi / 6), math.sin(math.pi / 6)

    def params(k: int):
        t = 15.0 * k / 499.0
        eps = 5e-3 * k / 499.0
        return t * dirx, t * diry, eps

    def frame(k: int) -> np.ndarray:
        tu, tv, eps = params(k)

        def inverse_map(x, y):
            return c[0] + (x - c[0] - tu) / (1.0 + eps), c[1] + (y - c[1] - tv) / (1.0 + eps)

        return warp_target(ref, inverse_map, np.uint16, 4095.0)

    def truth(k: int, X, Y):
        tu, tv, eps = params(k)
        X = np.asarray(X, dtype=np.float64)
        Y = np.asarray(Y, dtype=np.float64)
        one = np.ones_like(X)
        return tu + eps * (X - c[0]), tv + eps * (Y - c[1]), eps * one, 0 * one, 0 * one, eps * one

    pois = grid_queue(size, size, max(radii), grid)
    r = np.array([radii[i % len(radii)] for i in range(len(pois))], dtype=np.float32)
    pois["subset_radius"][:, 0] = r
    pois["subset_radius"][:, 1] = r
    meta = {"seed": seed, "size": size, "n_dots": n_dots, "frames": frames, "grid": grid, "margin": max(radii) + 2,
            "radii": list(radii), "radius_rule": "radii[i % 3]", "direction_deg": 30.0,
            "translation": "15 k / 499 px", "strain": "5e-3 k / 499 (ux = vy)", "noise_sigma": 24.0,
            "background": 640, "peak": 2880, "bit_depth": 12, "resample": "keys-cubic a=-0.5"}
    return Sequence(f"config2_{size}", ref, pois, frames, frame, truth, meta)


_POOL_SEQ = None


def _config2_frame(job):
    """Frame k of config2(size, frames) (process-pool worker)."""
    global _POOL_SEQ
    size, frames, k = job
    if _POOL_SEQ is None or _POOL_SEQ.meta.get("size") != size or _POOL_SEQ.frames != frames:
        _POOL_SEQ = config2(size=size, frames=frames)
    return _POOL_SEQ.frame(k)


def config2_frames(size: int = 1024, frames: int = 500, ks=None, workers: int = 16) -> list:
    """Frames ks (default all) of config2(size, frames), rendered by a pool of
    spawned worker processes (they use no GPU)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    ks = list(range(frames)) if ks is None else list(ks)
    workers = max(1, min(workers, os.cpu_count() or 1, len(ks)))
    with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_config2_frame, [(size, frames, k) for k in ks], chunksize=8))


def config4(size: int = 4096, grid: int = 4, radius: int = 20, seed: int = 20264, jump: float = 1.5,
            angle_deg: float = 30.0, slope: float = 1e-3) -> Scene:
    """Large field with a discontinuity: size^2 u16 speckle (16-bit range,
    background 10280, peak 46260, noise sigma 2 grey levels), a straight crack
    through the image centre at angle_deg, u = slope * X + jump on the crack's
    positive side (v = 0).  POIs on a `grid` px grid, subset radius `radius`,
    margin radius + 2; guess 0.  A target pixel in the opening (no preimage)
    takes the mean of its two one-sided preimages."""
    rng = np.random.default_rng(seed)
    n_dots = int(round(6000 * size * size / (512 * 512)))
    cx, cy, s = dot_field(size, size, n_dots, rng)
    ref = quantize(render_dots(cx, cy, s, size, size, 10280.0, 35980.0), rng, 2.0, np.uint16, 65535.0)
    c = ((size - 1) / 2.0, (size - 1) / 2.0)
    nx, ny = -math.sin(math.radians(angle_deg)), math.cos(math.radians(angle_deg))

    def side(X, Y):
        return (X - c[0]) * nx + (Y - c[1]) * ny

    def inverse_map(x, y):
        xa = x / (1.0 + slope)             # preimage on the negative side
        xb = (x - jump) / (1.0 + slope)    # on the positive side
        on_b = side(xb, y) > 0
        on_a = side(xa, y) <= 0
        X = np.where(on_b & ~on_a, xb, np.where(on_a & ~on_b, xa, 0.5 * (xa + xb)))
        return X, y

    tar = warp_target(ref, inverse_map, np.uint16, 65535.0)

    def truth(X, Y):
        X = np.asarray(X, dtype=np.float64)
        Y = np.asarray(Y, dtype=np.float64)
        u = slope * X + jump * (side(X, Y) > 0)
        z = np.zeros_like(X)
        return u, z, slope + z, z, z, z

    pois = grid_queue(size, size, radius, grid)
    meta = {"seed": seed, "n_dots": n_dots, "grid": grid, "margin": radius + 2, "bit_depth": 16,
            "noise_sigma": 2.0, "crack": {"angle_deg": angle_deg, "through": "centre", "jump_u_px": jump},
            "smooth_field": f"u = {slope} * x", "resample": "keys-cubic a=-0.5"}
    sc = Scene(f"config4_{size}", ref, tar, pois, radius, truth, meta)
    sc.crack_distance = lambda X, Y: np.abs(side(np.asarray(X, np.float64), np.asarray(Y, np.float64)))
    return sc


def lcg_pair(width: int = 192, height: int = 128, seed: int = 12345):
    """The reference's FFTCC known-answer images
    (smoke/library_boundary_probe.cpp:41-47): LCG pixels, target = reference
    moved by (+2, -1) with fill 127."""
    state = np.uint32(seed)
    ref = np.empty((height, width), dtype=np.float32)
    a, c = np.uint32(1664525), np.uint32(1013904223)
    with np.errstate(over="ignore"):
        for y in range(height):
            for x in range(width):
                state = np.uint32(a * state + c)
                ref[y, x] = float(state >> np.uint32(24))
    tar = np.full_like(ref, 127.0)
    tar[: height - 1, 2:] = ref[1:, : width - 2]
    pois = np.zeros(0, dtype=POI_DTYPE)
    pts = [(x, y) for y in range(32, 96, 16) for x in range(32, 160, 16)]
    from .opencorr import poi_queue
    pois = poi_queue(pts)
    return ref, tar, pois
