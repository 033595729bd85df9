"""ctypes binding of libfasticgn.so, the C ABI declared in include/fasticgn.h.

The library is built in-tree (``fasticgn_amd/libfasticgn.so``).  There is no
fallback: if it is missing or does not load, every entry point of this package
raises, so no result can silently come from anywhere but the gfx950 kernels.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libfasticgn.so"

FI_OK = 0
FI_E_INVALID_ARGUMENT = -1
FI_E_LOGIC = -2
FI_E_OUT_OF_RANGE = -3
FI_E_DEVICE = -4
FI_E_UNSUPPORTED = -5
FI_E_NO_MEMORY = -6

FI_PIXEL_U8, FI_PIXEL_U16, FI_PIXEL_F32 = 0, 1, 2
FI_DEPTH_NATIVE, FI_DEPTH12_RIGHT, FI_DEPTH12_LEFT, FI_DEPTH12_AUTO, FI_DEPTH_AUTO, FI_DEPTH16 = 0, 1, 2, 3, 4, 5
FI_NORMALIZE_NONE, FI_NORMALIZE_TO8, FI_NORMALIZE_TO16, FI_NORMALIZE_UNIT = 0x00, 0x10, 0x20, 0x30
FI_POLICY_SKIP_FFTCC = 1
FI_POLICY_POI_FALLBACK = 2
FI_POLICY_DEFAULT = FI_POLICY_SKIP_FFTCC | FI_POLICY_POI_FALLBACK
FI_SAMPLING_AUTO, FI_SAMPLING_TABLE, FI_SAMPLING_GROUP, FI_SAMPLING_PAIR = 0, 1, 2, 3
# FI_SEED_SAMPLE_*: the POIs the learned seeder's network seeds under the consensus
SEED_SAMPLES = {"all": 0, "checker": 1, "quarter": 2}
SAMPLING_MODES = {"auto": FI_SAMPLING_AUTO, "table": FI_SAMPLING_TABLE, "group": FI_SAMPLING_GROUP,
                  "pair": FI_SAMPLING_PAIR}
FI_POLICY_LEARNED_SEED = 4
FI_EXPORT_REF_F32, FI_EXPORT_GRAD_X, FI_EXPORT_GRAD_Y, FI_EXPORT_TAR_F32, FI_EXPORT_BSPLINE, FI_EXPORT_PREP = range(6)
# PoiPrep (fi_params.h), the records FI_EXPORT_PREP copies out: 176 bytes
PREP_DTYPE = np.dtype([("mean", "<f4"), ("norm", "<f4"), ("status", "<i4"), ("rx", "<i4"),
                       ("hinv", "<f4", (36,)), ("ref_sum", "<f4"), ("ry", "<i4"), ("pad", "<f4", (2,))])
assert PREP_DTYPE.itemsize == 176

# opencorr::POI2D, src/oc_poi.h:102-108: 25 float32 = 100 bytes
POI_DTYPE = np.dtype([
    ("x", "<f4"), ("y", "<f4"),
    ("deformation", "<f4", (12,)),
    ("result", "<f4", (6,)),
    ("strain", "<f4", (3,)),
    ("subset_radius", "<f4", (2,)),
])
assert POI_DTYPE.itemsize == 100


class FrameStats(ctypes.Structure):
    """fi_frame_stats (include/fasticgn.h)."""
    _fields_ = [
        ("frame_index", ctypes.c_int32),
        ("skip_fftcc_used", ctypes.c_int32),
        ("fallback_used", ctypes.c_int32),
        ("poi_fallback_count", ctypes.c_int32),
        ("poi_fallback_recovered", ctypes.c_int32),
        ("iteration_max", ctypes.c_int32),
        ("bad_ratio", ctypes.c_float),
        ("pre_repair_bad_ratio", ctypes.c_float),
        ("baseline_bad_ratio", ctypes.c_float),
        ("fallback_threshold", ctypes.c_float),
        ("valid_ratio", ctypes.c_float),
        ("iteration_mean", ctypes.c_float),
        ("valid_count", ctypes.c_int64),
        ("iteration_sum", ctypes.c_int64),
        ("fftcc_pois", ctypes.c_int64),
        ("device_ms", ctypes.c_double),
        ("bspline_ms", ctypes.c_double),
        ("fftcc_ms", ctypes.c_double),
        ("icgn_ms", ctypes.c_double),
        ("seed_pois", ctypes.c_int64),
        ("seed_fallback_count", ctypes.c_int64),
        ("seed_ms", ctypes.c_double),
        ("temporal_seed_count", ctypes.c_int32),
        ("temporal_seed_used", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class IngestInfo(ctypes.Structure):
    """fi_ingest_info (include/fasticgn.h)."""
    _fields_ = [("effective_bits", ctypes.c_int32), ("left_aligned", ctypes.c_int32),
                ("source_max", ctypes.c_double), ("scale", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class PrepareStats(ctypes.Structure):
    """fi_prepare_stats (include/fasticgn.h)."""
    _fields_ = [(name, ctypes.c_double) for name in
                ("total_ms", "alloc_ms", "reference_ms", "fftcc_ms", "hessian_ms")]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class TuneEntry(ctypes.Structure):
    """fi_tune_entry (include/fasticgn.h)."""
    _fields_ = [
        ("radius", ctypes.c_int32),
        ("waves", ctypes.c_int32),
        ("rows_per_band", ctypes.c_int32),
        ("sampling", ctypes.c_int32),
        ("persistent", ctypes.c_int32),
        ("chosen", ctypes.c_int32),
        ("ms", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class LogicError(RuntimeError):
    """std::logic_error of the reference."""


class DeviceError(RuntimeError):
    """A HIP runtime failure."""


_EXC = {
    FI_E_INVALID_ARGUMENT: ValueError,
    FI_E_LOGIC: LogicError,
    FI_E_OUT_OF_RANGE: IndexError,
    FI_E_DEVICE: DeviceError,
    FI_E_UNSUPPORTED: NotImplementedError,
    FI_E_NO_MEMORY: MemoryError,
}

_P = ctypes.c_void_p
_I = ctypes.c_int
_S = ctypes.c_size_t
_F = ctypes.c_float
_U = ctypes.c_uint

# name -> argtypes; every function returns int
SIGNATURES = {
    "fi_version": [ctypes.POINTER(_I)] * 3,
    "fi_device_count": [ctypes.POINTER(_I)],
    "fi_device_synchronize": [_I],
    "fi_last_error": [ctypes.c_char_p, _S],
    "fi_dft_twiddles": [_I, _P],
    "fi_image_resolve": [_P, _I, _I, _S, _I, _I, ctypes.POINTER(IngestInfo)],
    "fi_conv_threshold2": [_F, ctypes.POINTER(ctypes.c_double)],
    "fi_fftcc_create": [_I, _I, _I, ctypes.POINTER(_P)],
    "fi_fftcc_destroy": [_P],
    "fi_fftcc_set_ref": [_P, _P, _I, _I, _S, _I, _I],
    "fi_fftcc_set_tar": [_P, _P, _I, _I, _S, _I, _I],
    "fi_fftcc_prepare_ref_clear": [_P],
    "fi_fftcc_prepare_ref": [_P, _P, _S],
    "fi_fftcc_has_prepared_ref": [_P, ctypes.POINTER(_I)],
    "fi_fftcc_compute": [_P, _P, _S],
    "fi_fftcc_compute_indexed": [_P, _P, _S, _P],
    "fi_icgn_create": [_I, _I, _F, _F, _I, ctypes.POINTER(_P)],
    "fi_icgn_destroy": [_P],
    "fi_icgn_set_iteration": [_P, _F, _F],
    "fi_icgn_set_self_adaptive": [_P, _I],
    "fi_icgn_set_subset": [_P, _I, _I],
    "fi_icgn_set_shape": [_P, _I, _I],
    "fi_icgn_set_sampling": [_P, _I],
    "fi_icgn_set_ref": [_P, _P, _I, _I, _S, _I, _I],
    "fi_icgn_set_tar": [_P, _P, _I, _I, _S, _I, _I],
    "fi_icgn_prepare_ref": [_P],
    "fi_icgn_prepare_tar": [_P],
    "fi_icgn_prepare": [_P],
    "fi_icgn_compute": [_P, _P, _S],
    "fi_icgn_compute_offset": [_P, _P, _S, _P],
    "fi_icgn_export": [_P, _I, _P, _S],
    "fi_icgn_compute_noalloc": [_P, _P, _S],
    "fi_frame_create": [_I, _F, _F, _I, ctypes.POINTER(_P)],
    "fi_frame_destroy": [_P],
    "fi_frame_set_ref": [_P, _P, _I, _I, _S, _I, _I],
    "fi_frame_set_pois": [_P, _P, _S],
    "fi_frame_upload_target": [_P, _P, _I, _I, _S, _I, _I],
    "fi_frame_run": [_P, _U, ctypes.POINTER(FrameStats)],
    "fi_frame_run_async": [_P, _U],
    "fi_frame_wait": [_P, ctypes.POINTER(FrameStats)],
    "fi_frame_download": [_P, _P, _S],
    "fi_frame_reset_sequence": [_P],
    "fi_frame_set_shape": [_P, _I, _I],
    "fi_frame_set_sampling": [_P, _I],
    "fi_frame_set_seeder": [_P, _P, _F],
    "fi_frame_set_seed_passes": [_P, _I],
    "fi_frame_set_seed_consensus": [_P, _I, _F, _I],
    "fi_frame_set_self_adaptive": [_P, _I],
    "fi_frame_set_graph": [_P, _I],
    "fi_frame_set_persistent": [_P, _I],
    "fi_frame_download_redo": [_P, _P, _S],
    "fi_frame_set_fused": [_P, _I],
    "fi_frame_autotune": [_P, _P, _S, ctypes.POINTER(_S)],
    "fi_frame_records": [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(_S)],
    "fi_frame_set_records": [_P, _I],
    "fi_frame_prepare": [_P, ctypes.POINTER(PrepareStats)],
    "fi_frame_set_temporal_extrapolation": [_P, _I, _F],
    "fi_seeder_create": [_I, ctypes.POINTER(_P)],
    "fi_seeder_destroy": [_P],
    "fi_seeder_param_count": [ctypes.POINTER(_S)],
    "fi_seeder_set_params": [_P, _P, _S],
    "fi_seeder_get_params": [_P, _P, _S],
    "fi_seeder_set_training_data": [_P, _P, _P, _S],
    "fi_seeder_train_epoch": [_P, _I, _F, ctypes.c_uint64, ctypes.POINTER(_F)],
    "fi_seeder_forward": [_P, _P, _S, _P],
    "fi_seeder_gradient": [_P, _P, _P, _S, _P, ctypes.POINTER(_F)],
    "fi_seeder_seed": [_P, _P, _P, _I, _I, _P, _S, _P],
    "fi_seeder_seed_consensus": [_P, _P, _P, _I, _I, _P, _S, _P, _F, _F, _I, _I],
    "fi_format_raw": [_P, _S, _P, _S, ctypes.POINTER(_S)],
    "fi_write_raw": [_P, _S, ctypes.c_char_p, ctypes.POINTER(_S)],
}

_lib = None


def load() -> ctypes.CDLL:
    """Load libfasticgn.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("FASTICGN_LIB", LIB_PATH))
    if not path.exists():
        raise ImportError(f"{path} is not built; build it with make -C fasticgn_amd/csrc")
    lib = ctypes.CDLL(str(path))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _lib = lib
    return lib


def last_error() -> str:
    lib = load()
    buf = ctypes.create_string_buffer(1024)
    lib.fi_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int) -> None:
    if status != FI_OK:
        raise _EXC.get(status, RuntimeError)(last_error())


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def device_count() -> int:
    n = ctypes.c_int(0)
    status = load().fi_device_count(ctypes.byref(n))
    return n.value if status == FI_OK else 0


def synchronize(device: int = 0) -> None:
    call("fi_device_synchronize", int(device))


def ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)
