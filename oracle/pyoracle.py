"""ctypes wrapper of liboracle.so, the CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product (fasticgn_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

POI_DTYPE = np.dtype([
    ("x", "<f4"), ("y", "<f4"),
    ("deformation", "<f4", (12,)),
    ("result", "<f4", (6,)),
    ("strain", "<f4", (3,)),
    ("subset_radius", "<f4", (2,)),
])

_lib = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_S = ctypes.c_size_t
_F = ctypes.c_float
_D = ctypes.c_double


def build() -> None:
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        sig = {
            "or_image_to_float": ([_P, _I, _I, _I, _S, _I, _P], _I),
            "or_gradient_f32": ([_P, _I, _I, _P, _P, _I], None),
            "or_gradient_f64": ([_P, _I, _I, _P, _P, _I], None),
            "or_bspline_prepare_f32": ([_P, _I, _I, _P, _I], None),
            "or_bspline_prepare_f64": ([_P, _I, _I, _P, _I], None),
            "or_bspline_eval_f32": ([_P, _I, _I, _F, _F], _F),
            "or_bspline_eval_f64": ([_P, _I, _I, _D, _D], _D),
            "or_dft_twiddles": ([_I, _P], None),
            "or_fftcc_correlate": ([_P, _P, _I, _P], None),
            "or_fftcc_set_serial_sums": ([_I], None),
            "or_fftcc_compute": ([_P, _P, _I, _I, _I, _P, _S, _I], None),
            "or_fftcc_prepare_ref": ([_P, _I, _I, _I, _P, _S, _P, _P, _P, _P, _P, _I], None),
            "or_fftcc_compute_prepared": ([_P, _I, _I, _I, _P, _P, _P, _P, _S, _P, _I], None),
            "or_icgn_compute_f32": ([_P, _I, _I, _P, _P, _P, _I, _I, _F, _F, _I, _P, _S, _I], None),
            "or_icgn_compute_f64": ([_P, _I, _I, _P, _P, _P, _I, _I, _D, _D, _I, _P, _S, _I], None),
            "or_icgn_compute_offset_f32": ([_P, _I, _I, _P, _P, _P, _I, _I, _F, _F, _I, _P, _S, _P, _I], None),
            "or_icgn_compute_offset_f64": ([_P, _I, _I, _P, _P, _P, _I, _I, _D, _D, _I, _P, _S, _P, _I], None),
            "or_inverse6_f32": ([_P, _P], None),
            "or_inverse6_f64": ([_P, _P], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def threads() -> int:
    return int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))


DEPTH_MODES = {"native": 0, "right": 1, "left": 2, "auto12": 3, "auto": 4, "16": 5}
NORMALIZE_MODES = {"none": 0, "to8": 1, "to16": 2, "unit": 3}


def as_f32(img: np.ndarray, depth12: str = "native", normalize: str = "none") -> np.ndarray:
    """loadImage2DMat on decoded pixels: depth12 'native' (as stored),
    'right' / 'left' (12-bit), 'auto12' (12-bit, alignment inferred) or 'auto'
    (depth inferred); normalize 'none', 'to8', 'to16' or 'unit'."""
    a = np.ascontiguousarray(img)
    pixel = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}[a.dtype]
    mode = DEPTH_MODES[depth12] | (NORMALIZE_MODES[normalize] << 4)
    out = np.empty(a.shape, dtype=np.float32)
    rc = lib().or_image_to_float(_p(a), pixel, a.shape[1], a.shape[0], a.strides[0], mode, _p(out))
    if rc == -1:
        raise ValueError("12-bit right-aligned input has values above 4095")
    if rc != 0:
        raise ValueError(f"input mode {depth12}/{normalize} does not apply to {a.dtype} pixels")
    return out


def gradient(img: np.ndarray, f64: bool = False):
    img = np.ascontiguousarray(img, dtype=np.float32)
    h, w = img.shape
    dt = np.float64 if f64 else np.float32
    gx = np.empty((h, w), dtype=dt)
    gy = np.empty((h, w), dtype=dt)
    fn = lib().or_gradient_f64 if f64 else lib().or_gradient_f32
    fn(_p(img), w, h, _p(gx), _p(gy), threads())
    return gx, gy


def bspline(img: np.ndarray, f64: bool = False) -> np.ndarray:
    img = np.ascontiguousarray(img, dtype=np.float32)
    h, w = img.shape
    coef = np.empty((h, w, 16), dtype=np.float64 if f64 else np.float32)
    fn = lib().or_bspline_prepare_f64 if f64 else lib().or_bspline_prepare_f32
    fn(_p(img), w, h, _p(coef), threads())
    return coef


def bspline_eval(coef: np.ndarray, x: float, y: float) -> float:
    h, w = coef.shape[:2]
    if coef.dtype == np.float64:
        return lib().or_bspline_eval_f64(_p(coef), w, h, x, y)
    return lib().or_bspline_eval_f32(_p(coef), w, h, x, y)


def twiddles(n: int) -> np.ndarray:
    out = np.empty(2 * n, dtype=np.float32)
    lib().or_dft_twiddles(n, _p(out))
    return out


def correlate(ref_win: np.ndarray, tar_win: np.ndarray) -> np.ndarray:
    r = np.ascontiguousarray(ref_win, dtype=np.float32)
    t = np.ascontiguousarray(tar_win, dtype=np.float32)
    n = r.shape[0]
    out = np.empty((n, n), dtype=np.float32)
    lib().or_fftcc_correlate(_p(r), _p(t), n, _p(out))
    return out


def fftcc(ref: np.ndarray, tar: np.ndarray, radius: int, queue: np.ndarray, serial_sums: bool = False) -> None:
    """FFTCC2D::compute(POI2D*) over the queue (uncached), in place.
    serial_sums: window mean and norm in the reference's serial order instead
    of the device's row partials."""
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    tar = np.ascontiguousarray(tar, dtype=np.float32)
    assert queue.dtype == POI_DTYPE and queue.flags.c_contiguous
    h, w = ref.shape
    lib().or_fftcc_set_serial_sums(int(serial_sums))
    try:
        lib().or_fftcc_compute(_p(ref), _p(tar), w, h, radius, _p(queue), len(queue), threads())
    finally:
        lib().or_fftcc_set_serial_sums(0)


class PreparedRef:
    """FFTCC2D::prepareRef(queue) cache."""

    def __init__(self, ref: np.ndarray, radius: int, layout: np.ndarray):
        ref = np.ascontiguousarray(ref, dtype=np.float32)
        h, w = ref.shape
        n = len(layout)
        nh = 2 * radius * (radius + 1)
        self.radius, self.w, self.h = radius, w, h
        self.spectra = np.zeros((max(n, 1), nh, 2), dtype=np.float32)
        self.norm = np.zeros(max(n, 1), dtype=np.float32)
        self.valid = np.zeros(max(n, 1), dtype=np.int32)
        self.xs = np.zeros(max(n, 1), dtype=np.int32)
        self.ys = np.zeros(max(n, 1), dtype=np.int32)
        lay = np.ascontiguousarray(layout)
        lib().or_fftcc_prepare_ref(_p(ref), w, h, radius, _p(lay), n, _p(self.spectra),
                                   _p(self.norm), _p(self.valid), _p(self.xs), _p(self.ys), threads())

    def compute(self, tar: np.ndarray, queue: np.ndarray, idx=None) -> None:
        tar = np.ascontiguousarray(tar, dtype=np.float32)
        ip = None
        if idx is not None:
            idx = np.ascontiguousarray(idx, dtype=np.uint64)
            ip = _p(idx)
        lib().or_fftcc_compute_prepared(_p(tar), self.w, self.h, self.radius, _p(self.spectra),
                                        _p(self.norm), _p(self.valid), _p(queue), len(queue),
                                        ip, threads())


def icgn(ref: np.ndarray, tar: np.ndarray, radius, queue: np.ndarray, conv: float = 0.001,
         stop: float = 10.0, f64: bool = False, self_adaptive: bool = False,
         prepared=None, center_offsets: np.ndarray | None = None) -> None:
    """ICGN2D1 prepare() + compute(queue), in place.  radius: int or (rx, ry).
    center_offsets (n x 2, optional): compute(queue, center_offset_queue)."""
    rx, ry = (radius, radius) if np.isscalar(radius) else radius
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    if prepared is None:
        gx, gy = gradient(ref, f64)
        coef = bspline(tar, f64)
    else:
        gx, gy, coef = prepared
    h, w = ref.shape
    assert queue.dtype == POI_DTYPE and queue.flags.c_contiguous
    if center_offsets is None:
        fn = lib().or_icgn_compute_f64 if f64 else lib().or_icgn_compute_f32
        fn(_p(ref), w, h, _p(gx), _p(gy), _p(coef), int(rx), int(ry), conv, stop,
           int(self_adaptive), _p(queue), len(queue), threads())
    else:
        off = np.ascontiguousarray(center_offsets, dtype=np.float32)
        assert off.shape == (len(queue), 2)
        fn = lib().or_icgn_compute_offset_f64 if f64 else lib().or_icgn_compute_offset_f32
        fn(_p(ref), w, h, _p(gx), _p(gy), _p(coef), int(rx), int(ry), conv, stop,
           int(self_adaptive), _p(queue), len(queue), _p(off), threads())


def inverse6(a: np.ndarray, f64: bool = False) -> np.ndarray:
    dt = np.float64 if f64 else np.float32
    a = np.ascontiguousarray(a, dtype=dt)
    out = np.empty((6, 6), dtype=dt)
    (lib().or_inverse6_f64 if f64 else lib().or_inverse6_f32)(_p(a), _p(out))
    return out
