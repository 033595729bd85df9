"""Raw output tree of the multi-frame driver: one ``frame_NNNNNN.bin`` per
target (36-byte little-endian records, fi_write_raw) and the manifest
``manifest.opencorr-raw.json`` in the ``opencorr-gpu-raw-v1`` format
(writeRawManifest, opencorr_cli_cuda_mvp.cpp:2615-2732; field list and
per-frame QA keys as there, docs/output-format.md), so that the reference's
reader (pipeline/fasticgn_io.py:91-316: read_manifest, iter_frames,
read_raw_frame) loads it unchanged.

    with RawWriter(out_dir, ref_path="ref.tif", subset_radius=15, grid_space=5) as w:
        for k, tar in enumerate(targets):
            res, qa = drv.process(Image2D(tar))
            w.write_frame(res, qa, tar_path=f"tar_{k}.tif")

The manifest is rewritten atomically (temporary file + rename) after every
frame, so an interrupted run leaves a manifest that lists exactly the frames
whose files are complete.  With an output session (fasticgn_amd/ledger.py)
the frames and the manifest are published through its intent ledger.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time
from pathlib import Path

import numpy as np

from . import _native as nat
from .opencorr import _check_queue

RAW_FORMAT = "opencorr-gpu-raw-v1"
RAW_MANIFEST_NAME = "manifest.opencorr-raw.json"
RECORD_BYTES = 36
RAW_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("u", "<f4"), ("v", "<f4"), ("ux", "<f4"), ("uy", "<f4"),
                      ("vx", "<f4"), ("vy", "<f4"), ("zncc", "<f4")])
FIELDS = [{"name": n, "dtype": "int32" if n in ("x", "y") else "float32", "offset": 4 * i}
          for i, n in enumerate(RAW_DTYPE.names)]
CSV_FIELD_ORDER = ["x", "y", "u", "v", "zncc", "uy", "vy", "ux", "vx"]


def write_raw(queue: np.ndarray, path) -> int:
    """One frame file through fi_write_raw; returns the bytes written."""
    q = _check_queue(queue)
    n = ctypes.c_size_t(0)
    nat.call("fi_write_raw", nat.ptr(q), len(q), os.fsencode(str(path)), ctypes.byref(n))
    return n.value


def format_raw(queue: np.ndarray) -> np.ndarray:
    """The 36-byte records of a POI queue through fi_format_raw (RAW_DTYPE)."""
    q = _check_queue(queue)
    out = np.zeros(len(q), dtype=RAW_DTYPE)
    n = ctypes.c_size_t(0)
    if len(q):
        nat.call("fi_format_raw", nat.ptr(q), len(q), nat.ptr(out), out.nbytes, ctypes.byref(n))
    return out


def _nullable(v, kind=float):
    if v is None:
        return None
    if kind is float:
        v = float(v)
        return None if not math.isfinite(v) else v
    return kind(v)


class RawWriter:
    """Writes a raw output tree for one reference and a sequence of targets."""

    def __init__(self, output_dir, ref_path: str = "", tar_list_path: str = "", subset_radius: int = 0,
                 grid_space: int = 0, conv_threshold: float = 0.001, max_iteration: int = 10,
                 policy: int | None = None, depth: str = "native", learned_seed: bool = False,
                 temporal_extrapolation: bool = False, session=None):
        from .driver import DEFAULT_POLICY, POI_FALLBACK, SKIP_FFTCC
        self.session = session
        self.dir = Path(output_dir)
        self.dir.mkdir(parents=True, exist_ok=True)
        policy = DEFAULT_POLICY if policy is None else policy
        self.head = {
            "format": RAW_FORMAT,
            "endianness": "little",
            "record_bytes": RECORD_BYTES,
            "ref_path": str(ref_path),
            "tar_list_path": str(tar_list_path),
            "output_dir": str(output_dir),
            "subset_radius": int(subset_radius),
            "grid_space": int(grid_space),
            "conv_threshold": float(conv_threshold),
            "max_iteration": int(max_iteration),
            "durability": "rename",
            "engine": {
                "implementation": "fasticgn_amd",
                "icgn_backend": "hip-gfx950",
                "icgn_kernel": "block-per-poi-column-bands",
                "oob_policy": "geometric-v1",
                "fftcc_backend": "hip-dft",
                "skip_fftcc": bool(policy & SKIP_FFTCC),
                "poi_fallback": bool(policy & POI_FALLBACK),
                "temporal_extrapolation": bool(temporal_extrapolation),
                "frame_fallback_policy": "per-poi-merge-v1",
                "learned_seed": bool(learned_seed),
            },
            "input_image": {"depth": depth, "align": "none", "normalize": "none"},
            "fields": FIELDS,
            "csv_field_order": CSV_FIELD_ORDER,
        }
        self.frames: list[dict] = []

    def write_frame(self, results: np.ndarray, qa: dict | None = None, tar_path: str = "",
                    output_name: str | None = None) -> dict:
        """Write the next frame file and its manifest entry; qa is the frame
        driver's stats dict (fi_frame_stats) or None."""
        k = len(self.frames)
        name = output_name or f"frame_{k:06d}.bin"
        t0 = time.perf_counter()
        published = None
        if self.session is not None:
            from .ledger import OutputSpec
            rec = results if results.dtype == RAW_DTYPE else format_raw(results)
            payload = np.ascontiguousarray(rec).tobytes()
            spec = OutputSpec(key="frame/" + name, path=self.dir / name, rows=len(rec), record_bytes=RECORD_BYTES)
            published = self.session.publish(spec, payload)
            nbytes = len(payload)
        elif results.dtype == RAW_DTYPE:
            # records already packed (FrameDriver.records())
            path = self.dir / name
            tmp = path.with_suffix(path.suffix + ".tmp")
            tmp.write_bytes(np.ascontiguousarray(results).tobytes())
            os.replace(tmp, path)
            nbytes = results.nbytes
        else:
            nbytes = write_raw(results, self.dir / name)
        t1 = time.perf_counter()
        qa = qa or {}
        entry = {
            "tar_path": str(tar_path), "output_name": name, "file": name,
            "rows": nbytes // RECORD_BYTES, "bytes": nbytes, "format_sec": 0.0, "write_sec": t1 - t0,
            "bad_ratio": _nullable(qa.get("bad_ratio")),
            "baseline_bad_ratio": _nullable(qa.get("baseline_bad_ratio")),
            "fallback_threshold": _nullable(qa.get("fallback_threshold")),
            "fallback_used": _nullable(qa.get("fallback_used"), bool),
            "skip_fftcc_used": _nullable(qa.get("skip_fftcc_used"), bool),
            "valid_ratio": _nullable(qa.get("valid_ratio")),
            "iteration_mean": _nullable(qa.get("iteration_mean")),
            "iteration_max": _nullable(qa.get("iteration_max"), int),
            "poi_fallback_count": _nullable(qa.get("poi_fallback_count"), int),
            "poi_fallback_recovered": _nullable(qa.get("poi_fallback_recovered"), int),
            # seeds set by the two-frame extrapolation (cli:3063-3067, 675-678)
            "temporal_seed_used": _nullable(qa.get("temporal_seed_used"), bool),
            "temporal_seed_count": _nullable(qa.get("temporal_seed_count"), int),
        }
        self.frames.append(entry)
        self.flush()
        return {**entry, "ledger": published} if published else entry

    def flush(self) -> Path:
        path = self.dir / RAW_MANIFEST_NAME
        text = json.dumps({**self.head, "frames": self.frames}, indent=2) + "\n"
        if self.session is not None:
            self.session.publish_metadata(path, text.encode())
            return path
        tmp = path.with_suffix(".json.tmp")
        tmp.write_text(text)
        os.replace(tmp, path)
        return path

    def close(self) -> None:
        self.flush()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_raw_frame(path) -> np.ndarray:
    payload = Path(path).read_bytes()
    if len(payload) % RECORD_BYTES:
        raise ValueError(f"{path} holds {len(payload)} bytes, not a multiple of {RECORD_BYTES}")
    return np.frombuffer(payload, dtype=RAW_DTYPE)
