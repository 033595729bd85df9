"""The whole-frame fallback scenarios of the reference's gate
(smoke/run_frame_fallback_gate.py:8-57, cases at :94-104), rendered here with
numpy and this project's own seeds, and their six assertions.

Each sequence sums Gaussian speckles (sigma 1.6 px, 0.08 per px^2, amplitude
U[0.5, 1]) translated by the exact sub-pixel displacement of its frame, so
every POI's true displacement is known; frames are quantised to 8 bit around
128 with a gain of 40 / std(reference) and Gaussian noise sigma 2.

cascade  192x144 px, r = 7, grid 2: translation growing 0.5 px per frame from
         3.25 to 9.75 px, past what FFTCC from a zero guess captures at r = 7,
         with a 40x40 px occluder (foreign speckle) on frame 6.
jump     320x240 px, r = 15, grid 4: a 10 px jump between frames 0 and 1.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from fasticgn_amd.opencorr import grid_queue

GOOD = 0.8
SIGMA = 1.6
DENSITY = 0.08
NOISE = 2.0


@dataclass(frozen=True)
class Case:
    name: str
    width: int
    height: int
    radius: int
    grid: int
    truth: tuple
    occluded_frame: int | None
    occluder: tuple | None
    floor: float
    floor_from: int


CASES = {
    "cascade": Case("cascade", 192, 144, 7, 2, tuple((3.25 + 0.5 * k, -0.4 + 0.05 * k) for k in range(14)),
                    6, (76, 52, 40), 0.90, 13),
    "jump": Case("jump", 320, 240, 15, 4, ((0.25, 0.1), (10.25, 0.1), (10.75, 0.1), (11.25, 0.1)),
                 None, None, 0.99, 1),
}


def speckles(rng, x_lo, x_hi, y_lo, y_hi):
    count = int(round(DENSITY * (x_hi - x_lo) * (y_hi - y_lo)))
    return (rng.uniform(x_lo, x_hi, count), rng.uniform(y_lo, y_hi, count), rng.uniform(0.5, 1.0, count))


def render(width, height, spk, dx, dy):
    """Sum of the speckles translated by (dx, dy), each over the pixels within
    4 sigma of its centre along both axes (float64)."""
    sx, sy, amp = spk
    cx, cy = sx + dx, sy + dy
    reach = 4.0 * SIGMA
    k = int(math.ceil(reach)) + 1
    fx, fy = np.floor(cx).astype(np.int64), np.floor(cy).astype(np.int64)
    acc = np.zeros(width * height)
    s = -0.5 / (SIGMA * SIGMA)
    for oy in range(-k, k + 1):
        iy = fy + oy
        ey = iy - cy
        wy = amp * np.exp(s * ey * ey)
        oky = (iy >= 0) & (iy < height) & (np.abs(ey) <= reach)
        for ox in range(-k, k + 1):
            ix = fx + ox
            ex = ix - cx
            ok = oky & (ix >= 0) & (ix < width) & (np.abs(ex) <= reach)
            acc += np.bincount((iy * width + ix)[ok], (wy * np.exp(s * ex * ex))[ok], minlength=acc.size)
    return acc.reshape(height, width)


def quantize(img, mean, std, rng):
    gray = 128.0 + (40.0 / std) * (img - mean) + rng.normal(0.0, NOISE, img.shape)
    return np.clip(np.floor(gray + 0.5), 0, 255).astype(np.uint8)


def sequence(case: Case, seed: int = 20261204):
    """(reference, [targets]) as uint8 arrays."""
    w, h = case.width, case.height
    reach = 4.0 * SIGMA + 1.0
    us, vs = [t[0] for t in case.truth], [t[1] for t in case.truth]
    rng = np.random.default_rng(seed)
    spk = speckles(rng, -reach - max(0.0, max(us)), w + reach - min(0.0, min(us)),
                   -reach - max(0.0, max(vs)), h + reach - min(0.0, min(vs)))
    base = render(w, h, spk, 0.0, 0.0)
    mean, std = float(base.mean()), float(base.std())
    ref = quantize(base, mean, std, np.random.default_rng(seed + 100))
    foreign = None
    if case.occluder is not None:
        foreign = render(w, h, speckles(np.random.default_rng(seed + 1), -reach, w + reach, -reach, h + reach),
                         0.0, 0.0)
    tars = []
    for k, (u, v) in enumerate(case.truth):
        img = render(w, h, spk, u, v)
        if k == case.occluded_frame:
            ox, oy, side = case.occluder
            img[oy:oy + side, ox:ox + side] = foreign[oy:oy + side, ox:ox + side]
        tars.append(quantize(img, mean, std, np.random.default_rng(seed + 101 + k)))
    return ref, tars


def template(case: Case):
    return grid_queue(case.width, case.height, case.radius, case.grid)


def measurable(case: Case, q, u, v):
    """Target subset inside the interpolatable range at the true displacement,
    1 px slack (run_frame_fallback_gate.py:237-242)."""
    r = case.radius
    x, y = q["x"].astype(np.float64), q["y"].astype(np.float64)
    return (x + u - r >= 2.0) & (y + v - r >= 2.0) & (x + u + r < case.width - 3.0) & (y + v + r < case.height - 3.0)


def hits_occluder(case: Case, q, u, v):
    ox, oy, side = case.occluder
    r = case.radius
    x, y = q["x"].astype(np.float64), q["y"].astype(np.float64)
    return (x + u + r >= ox) & (x + u - r < ox + side) & (y + v + r >= oy) & (y + v - r < oy + side)


def valid(z):
    return np.isfinite(z) & (z >= 0)


def good(z):
    return np.isfinite(z) & (z >= GOOD)


def check_case(case: Case, frames, stats, scratch_frames):
    """The reference gate's assertions 1, 2, 3, 5 and 6 on one run of a case
    (assertion 4, independence from the thread count, is checked by the
    caller across two runs).  frames[k]: results of frame k; stats[k]: its QA
    (fallback_used, pre_repair_bad_ratio, bad_ratio); scratch_frames[k]: the
    from-scratch pass of frame k.  Returns (errors, per-frame report)."""
    errors, report = [], []
    for k, (q, st) in enumerate(zip(frames, stats)):
        u_t, v_t = case.truth[k]
        z = q["result"][:, 2]
        d = q["deformation"]
        meas = measurable(case, q, u_t, v_t)
        cover = float(valid(z[meas]).mean()) if meas.any() else float("nan")
        ok = valid(z)
        if k == case.occluded_frame:
            ok &= ~hits_occluder(case, q, u_t, v_t)
        err = np.maximum(np.abs(d[ok, 0] - u_t), np.abs(d[ok, 6] - v_t)).astype(np.float64)
        hiconf = int((good(z[ok]) & (err > 1.0)).sum())
        row = {"frame": k, "fallback": int(st["fallback_used"]), "trigger": float(st["pre_repair_bad_ratio"]),
               "after": float(st["bad_ratio"]), "coverage": cover,
               "median_err": float(np.median(err)) if len(err) else float("nan"),
               "within_0.1": float((err <= 0.1).mean()) if len(err) else float("nan"), "hiconf": hiconf}
        tag = f"{case.name} frame {k}"
        if st["fallback_used"]:
            if row["after"] > row["trigger"]:  # assertion 2
                errors.append(f"{tag}: fallback raised the bad ratio {row['trigger']} -> {row['after']}")
            sc = scratch_frames[k]["result"][:, 2]
            lost = good(sc) & ~good(z)  # assertion 3
            row["scratch_good_lost"] = int(lost.sum())
            if lost.any():
                errors.append(f"{tag}: {int(lost.sum())} POIs good in the from-scratch pass are not good, "
                              f"{int((lost & (z == -3)).sum())} of them -3")
        if k >= case.floor_from and not cover >= case.floor:  # assertion 5
            errors.append(f"{tag}: valid coverage {cover:.4f} < floor {case.floor}")
        if len(err):  # assertion 6
            if row["median_err"] > 0.02:
                errors.append(f"{tag}: median error {row['median_err']:.4f} px > 0.02")
            if row["within_0.1"] < 0.99:
                errors.append(f"{tag}: only {row['within_0.1']:.4f} of valid POIs within 0.1 px")
        if hiconf:
            errors.append(f"{tag}: {hiconf} POIs with zncc >= 0.8 and error > 1 px")
        report.append(row)
    if not any(st["fallback_used"] for st in stats):  # assertion 1
        errors.append(f"{case.name}: no frame took the whole-frame fallback; the case is void")
    return errors, report
