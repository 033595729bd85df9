"""Parity gate between two POI result sets (test infrastructure).

compare_frame   restates the reference's per-frame gate
                (smoke/check_cpu_gpu_parity.py:46-98): POIs keyed by integer
                (x, y); valid = finite u, v, zncc with zncc >= 0; on the POIs
                valid on both sides, per-field median <= 1e-3, nearest-rank
                p99 <= 1e-2 and max <= 0.5; at least 100 shared valid POIs,
                CPU-valid retention >= 0.95 and validity disagreement <= 0.05.
                `fields` widens the gated fields: the project gates u, v and
                the four displacement gradients ux, uy, vx, vy with the same
                thresholds.
compare         the project's gate for two result arrays of one POI queue:
                compare_frame over u, v, ux, uy, vx, vy, plus validity flags
                (valid / -3 / -4 / -5 / NaN) identical POI by POI; iteration
                count and convergence agreement on the shared valid POIs are
                reported, and gated when a bound is given.
check_dirs      the reference's frame-directory runner (:100-176): frames
                paired by index (frame_000000.csv, ...) or by a glob, one
                verdict per frame, a pooled shared-valid floor.  Run as
                `python tests/parity.py --cpu-dir A --gpu-dir B ...`.
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import struct
import sys
from pathlib import Path

import numpy as np

GRADIENTS = ("ux", "uy", "vx", "vy")
ALL_FIELDS = ("u", "v") + GRADIENTS
# POI2D deformation.p order: u ux uy uxx uxy uyy v vx vy vxx vxy vyy (src/oc_poi.h:27-33)
_P = {"u": 0, "ux": 1, "uy": 2, "v": 6, "vx": 7, "vy": 8}


def nearest_rank(values, fraction: float) -> float:
    """Nearest-rank quantile with no interpolation (check_cpu_gpu_parity.py:41-43)."""
    s = sorted(values)
    return s[max(0, math.ceil(len(s) * fraction) - 1)]


def row_valid(row: dict) -> bool:
    """check_cpu_gpu_parity.py:37-38: zncc >= 0 and u, v, zncc finite."""
    return row["zncc"] >= 0.0 and all(math.isfinite(row[f]) for f in ("u", "v", "zncc"))


def rows_from_pois(q: np.ndarray) -> dict:
    """POI array -> {(int x, int y): {u, v, ux, uy, vx, vy, zncc, iteration,
    convergence}}; a repeated location is an error, as in the reference's
    reader."""
    d = q["deformation"].astype(np.float64)
    r = q["result"].astype(np.float64)
    rows = {}
    for i in range(len(q)):
        key = (int(q["x"][i]), int(q["y"][i]))
        if key in rows:
            raise ValueError(f"duplicate POI {key}")
        row = {f: float(d[i, j]) for f, j in _P.items()}
        row["zncc"], row["iteration"], row["convergence"] = float(r[i, 2]), float(r[i, 3]), float(r[i, 4])
        rows[key] = row
    return rows


def compare_frame(cpu_rows: dict, gpu_rows: dict, *, median_threshold: float = 1e-3,
                  p99_threshold: float = 1e-2, max_threshold: float = 0.5,
                  min_shared_valid: int = 100, min_valid_retention: float = 0.95,
                  max_validity_disagreement: float = 0.05, fields=("u", "v")) -> dict:
    common = sorted(cpu_rows.keys() & gpu_rows.keys())
    n_cpu = n_gpu = 0
    lost, gained, shared = [], [], []
    for key in common:
        cv, gv = row_valid(cpu_rows[key]), row_valid(gpu_rows[key])
        n_cpu += cv
        n_gpu += gv
        if cv and not gv:
            lost.append(key)
        elif gv and not cv:
            gained.append(key)
        elif cv and gv:
            shared.append(key)
    retention = len(shared) / n_cpu if n_cpu else 0.0
    disagreement = (len(lost) + len(gained)) / len(common) if common else 1.0
    out = {"shared": len(common), "cpu_valid": n_cpu, "gpu_valid": n_gpu, "shared_valid": len(shared),
           "valid_retention": retention, "validity_disagreement": disagreement,
           "cpu_valid_gpu_invalid": len(lost), "gpu_valid_cpu_invalid": len(gained), "errors": []}
    if lost:
        xs, ys = [k[0] for k in lost], [k[1] for k in lost]
        out["lost_valid_bbox"] = [min(xs), min(ys), max(xs), max(ys)]
    if not common:
        out["errors"].append("no shared POIs")
    if len(shared) < min_shared_valid:
        out["errors"].append(f"shared valid POIs below per-frame minimum {min_shared_valid}")
    if retention < min_valid_retention:
        out["errors"].append(f"CPU-valid retention below {min_valid_retention:g}")
    if disagreement > max_validity_disagreement:
        out["errors"].append(f"validity disagreement exceeds {max_validity_disagreement:g}")
    if shared:
        over = {"median": False, "p99": False, "max": False}
        for f in tuple(fields) + ("zncc",):
            diff = [abs(cpu_rows[k][f] - gpu_rows[k][f]) for k in shared]
            stats = {"median": float(np.median(diff)), "p99": nearest_rank(diff, 0.99), "max": max(diff)}
            for stat, val in stats.items():
                out[f"{stat}_abs_d{f}"] = val
            if f == "zncc":
                continue
            out[f"max_d{f}_at"] = shared[int(np.argmax(diff))]
            # a NaN statistic (a non-finite gated field) fails as well
            over["median"] |= not stats["median"] <= median_threshold
            over["p99"] |= not stats["p99"] <= p99_threshold
            over["max"] |= not stats["max"] <= max_threshold
        names = "/".join(fields)
        if over["median"]:
            out["errors"].append(f"per-frame median {names} error exceeds {median_threshold:g}")
        if over["p99"]:
            out["errors"].append(f"per-frame p99 {names} error exceeds {p99_threshold:g}")
        if over["max"]:
            out["errors"].append(f"per-frame maximum {names} error exceeds {max_threshold:g}")
    out["pass"] = not out["errors"]
    return out


def finite_valid(q: np.ndarray) -> np.ndarray:
    """Validity of a POI array by the reference's rule."""
    z = q["result"][:, 2]
    d = q["deformation"]
    return (z >= 0) & np.isfinite(z) & np.isfinite(d[:, 0]) & np.isfinite(d[:, 6])


def status(q: np.ndarray) -> np.ndarray:
    """1 for a valid POI, else its zncc sentinel (-3 / -4 / -5 / other), -99 for NaN."""
    z = q["result"][:, 2].astype(np.float64)
    return np.where(finite_valid(q), 1.0, np.where(np.isfinite(z), z, -99.0))


def compare(cpu: np.ndarray, gpu: np.ndarray, *, median_threshold=1e-3, p99_threshold=1e-2,
            max_threshold=0.5, min_shared_valid=100, max_iteration_mismatch=None,
            fields=ALL_FIELDS) -> dict:
    """The project's gate (see the module docstring).  Keys median_du ...
    max_dv, median_dux ... name the gated statistics; iteration_mismatch is
    the number of shared valid POIs whose iteration counts differ and
    max_iteration_mismatch, when given, bounds its fraction."""
    assert len(cpu) == len(gpu)
    assert np.array_equal(cpu["x"], gpu["x"]) and np.array_equal(cpu["y"], gpu["y"])
    rc, rg = rows_from_pois(cpu), rows_from_pois(gpu)
    out = compare_frame(rc, rg, median_threshold=median_threshold, p99_threshold=p99_threshold,
                        max_threshold=max_threshold, min_shared_valid=min_shared_valid, fields=fields)
    out["n"] = len(cpu)
    for f in tuple(fields) + ("zncc",):
        for stat in ("median", "p99", "max"):
            if f"{stat}_abs_d{f}" in out:
                out[f"{stat}_d{f}"] = out[f"{stat}_abs_d{f}"]
    sc, sg = status(cpu), status(gpu)
    out["status_mismatch"] = int((sc != sg).sum())
    if out["status_mismatch"]:
        out["errors"].append("validity flags differ")
    both = finite_valid(cpu) & finite_valid(gpu)
    if both.any():
        itc, itg = cpu["result"][both, 3], gpu["result"][both, 3]
        out["iteration_mismatch"] = int((itc != itg).sum())
        out["iteration_mismatch_frac"] = out["iteration_mismatch"] / int(both.sum())
        out["max_diteration"] = float(np.max(np.abs(itc.astype(np.float64) - itg)))
        out["mean_iteration_cpu"], out["mean_iteration_gpu"] = float(itc.mean()), float(itg.mean())
        same_it = itc == itg
        if same_it.any():
            dc = np.abs(cpu["result"][both, 4].astype(np.float64) - gpu["result"][both, 4])[same_it]
            out["median_dconvergence"], out["max_dconvergence"] = float(np.median(dc)), float(dc.max())
        if max_iteration_mismatch is not None and out["iteration_mismatch_frac"] > max_iteration_mismatch:
            out["errors"].append(f"iteration counts differ on more than {max_iteration_mismatch:g} of the POIs")
    out["pass"] = not out["errors"]
    return out


# ------------------------------------------------------- frame directories --

RAW_RECORD = struct.Struct("<iifffffff")  # x y u v ux uy vx vy zncc (36 B)
DEFAULT_FRAMES = 10


def read_frame(path: Path) -> dict:
    """A frame file -> rows: CSV with x, y, u, v, zncc columns (gradients
    optional), or 36-byte raw records (.raw / .bin)."""
    rows = {}
    if path.suffix in (".raw", ".bin"):
        blob = path.read_bytes()
        if len(blob) % RAW_RECORD.size:
            raise ValueError(f"{path}: {len(blob)} bytes is not a whole number of records")
        for x, y, u, v, ux, uy, vx, vy, z in RAW_RECORD.iter_unpack(blob):
            if (x, y) in rows:
                raise ValueError(f"{path}: duplicate POI {(x, y)}")
            rows[(x, y)] = {"u": u, "v": v, "ux": ux, "uy": uy, "vx": vx, "vy": vy, "zncc": z}
        return rows
    with path.open("r", encoding="utf-8", newline="") as fh:
        for row in csv.DictReader(fh):
            key = (int(float(row["x"])), int(float(row["y"])))
            if key in rows:
                raise ValueError(f"{path}: duplicate POI {key}")
            rows[key] = {f: float(row[f]) for f in ("u", "v", "zncc")}
    return rows


def frame_names(cpu_dir: Path, gpu_dir: Path, frames=None, pattern=None) -> list:
    if pattern is None:
        return [f"frame_{i:06d}.csv" for i in range(DEFAULT_FRAMES if frames is None else frames)]
    names = sorted({p.name for d in (cpu_dir, gpu_dir) for p in d.glob(pattern) if p.is_file()})
    return names if frames is None else names[:frames]


def check_dirs(cpu_dir: Path, gpu_dir: Path, *, frames=None, pattern=None, min_shared_valid=100,
               min_frame_shared_valid=100, **thresholds) -> dict:
    results = []
    for name in frame_names(cpu_dir, gpu_dir, frames, pattern):
        try:
            res = compare_frame(read_frame(cpu_dir / name), read_frame(gpu_dir / name),
                                min_shared_valid=min_frame_shared_valid, **thresholds)
        except (OSError, ValueError, KeyError, TypeError) as ex:
            res = {"pass": False, "shared_valid": 0, "errors": [str(ex)]}
        res["frame"] = name
        results.append(res)
    pooled = sum(r["shared_valid"] for r in results)
    return {"schema_version": 1, "frames": results, "pooled_shared_valid": pooled,
            "pairing": "glob" if pattern else "index", "glob": pattern,
            "pass": bool(results) and pooled >= min_shared_valid and all(r["pass"] for r in results)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Per-frame CPU/GPU parity gate over two frame directories.")
    ap.add_argument("--cpu-dir", required=True, type=Path)
    ap.add_argument("--gpu-dir", required=True, type=Path)
    ap.add_argument("--frames", type=int, default=None)
    ap.add_argument("--glob", default=None, metavar="PATTERN")
    ap.add_argument("--median-threshold", type=float, default=1e-3)
    ap.add_argument("--p99-threshold", type=float, default=1e-2)
    ap.add_argument("--max-threshold", type=float, default=0.5)
    ap.add_argument("--min-shared-valid", type=int, default=100)
    ap.add_argument("--min-frame-shared-valid", type=int, default=100)
    ap.add_argument("--min-valid-retention", type=float, default=0.95)
    ap.add_argument("--max-validity-disagreement", type=float, default=0.05)
    ap.add_argument("--json-report", type=Path)
    a = ap.parse_args(argv)
    if (a.frames is not None and a.frames < 1) or min(a.min_shared_valid, a.min_frame_shared_valid) < 1:
        ap.error("frames and minimum shared-valid counts must be positive")
    for name in ("median_threshold", "p99_threshold", "max_threshold", "min_valid_retention",
                 "max_validity_disagreement"):
        val = getattr(a, name)
        if not math.isfinite(val) or val < 0:
            ap.error(f"{name} must be finite and nonnegative")
    if a.min_valid_retention > 1 or a.max_validity_disagreement > 1:
        ap.error("validity fractions must not exceed 1")
    rep = check_dirs(a.cpu_dir, a.gpu_dir, frames=a.frames, pattern=a.glob,
                     min_shared_valid=a.min_shared_valid, min_frame_shared_valid=a.min_frame_shared_valid,
                     median_threshold=a.median_threshold, p99_threshold=a.p99_threshold,
                     max_threshold=a.max_threshold, min_valid_retention=a.min_valid_retention,
                     max_validity_disagreement=a.max_validity_disagreement)
    for r in rep["frames"]:
        print(json.dumps(r, sort_keys=True, default=list))
    if a.json_report:
        a.json_report.write_text(json.dumps(rep, indent=2, default=list) + "\n")
    print(f"pooled_shared_valid={rep['pooled_shared_valid']} verdict={'PASS' if rep['pass'] else 'FAIL'}")
    return 0 if rep["pass"] else 1


if __name__ == "__main__":
    sys.exit(main())
