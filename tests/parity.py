"""Parity rules between two POI result sets, restating the reference's gate
smoke/check_cpu_gpu_parity.py:41-98 (nearest-rank p99, valid = finite zncc >= 0,
thresholds median 1e-3, p99 1e-2, max 0.5 px) plus the stricter rule of this
project: validity flags must match exactly."""
from __future__ import annotations

import math

import numpy as np


def nearest_rank(values, fraction: float) -> float:
    """smoke/check_cpu_gpu_parity.py:41-43"""
    s = sorted(values)
    return s[max(0, math.ceil(len(s) * fraction) - 1)]


def finite_valid(q: np.ndarray) -> np.ndarray:
    z = q["result"][:, 2]
    return (z >= 0) & np.isfinite(z) & np.isfinite(q["deformation"][:, 0]) & np.isfinite(q["deformation"][:, 6])


def compare(cpu: np.ndarray, gpu: np.ndarray, median_threshold=1e-3, p99_threshold=1e-2,
            max_threshold=0.5) -> dict:
    assert len(cpu) == len(gpu)
    assert np.array_equal(cpu["x"], gpu["x"]) and np.array_equal(cpu["y"], gpu["y"])
    cv, gv = finite_valid(cpu), finite_valid(gpu)
    both = cv & gv
    zc, zg = cpu["result"][:, 2], gpu["result"][:, 2]
    # status: valid, or the sentinel itself
    status_c = np.where(cv, 1.0, np.where(np.isfinite(zc), zc, -99.0))
    status_g = np.where(gv, 1.0, np.where(np.isfinite(zg), zg, -99.0))
    out = {"n": len(cpu), "cpu_valid": int(cv.sum()), "gpu_valid": int(gv.sum()),
           "shared_valid": int(both.sum()),
           "status_mismatch": int((status_c != status_g).sum()), "errors": []}
    if both.any():
        du = np.abs(cpu["deformation"][both, 0].astype(np.float64) - gpu["deformation"][both, 0])
        dv = np.abs(cpu["deformation"][both, 6].astype(np.float64) - gpu["deformation"][both, 6])
        for name, d in (("du", du), ("dv", dv)):
            out[f"median_{name}"] = float(np.median(d))
            out[f"p99_{name}"] = float(nearest_rank(d.tolist(), 0.99))
            out[f"max_{name}"] = float(d.max())
        if max(out["median_du"], out["median_dv"]) > median_threshold:
            out["errors"].append("median")
        if max(out["p99_du"], out["p99_dv"]) > p99_threshold:
            out["errors"].append("p99")
        if max(out["max_du"], out["max_dv"]) > max_threshold:
            out["errors"].append("max")
    if out["status_mismatch"]:
        out["errors"].append("validity flags differ")
    out["pass"] = not out["errors"]
    return out
