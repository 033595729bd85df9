"""The C ABI library loads and exports every symbol include/fasticgn.h
declares; argument validation that needs no device; the raw record format.
CPU only: no compute call reaches a GPU here."""
import ctypes
import re
import struct
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "fasticgn.h").read_text()
    return sorted(set(re.findall(r"^int\s+(fi_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from fasticgn_amd import _native
    lib = _native.load()
    names = declared_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.SIGNATURES), set(names) ^ set(_native.SIGNATURES)


def test_poi_layout_matches_reference():
    """opencorr::POI2D: x, y, p[12], r[6], e[3], subset_radius -> 100 bytes."""
    from fasticgn_amd import POI_DTYPE
    assert POI_DTYPE.itemsize == 100
    assert POI_DTYPE.fields["deformation"][1] == 8
    assert POI_DTYPE.fields["result"][1] == 56
    assert POI_DTYPE.fields["strain"][1] == 80
    assert POI_DTYPE.fields["subset_radius"][1] == 92


def test_version_and_errors_without_device():
    from fasticgn_amd import _native as nat
    lib = nat.load()
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.fi_version(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 0
    h = ctypes.c_void_p()
    # src/oc_fftcc.cpp:154-158: rectangular / non-positive / too large radii are invalid
    for rx, ry in [(7, 15), (15, 7), (0, 0), (-1, -1), (23171, 23171)]:
        assert lib.fi_fftcc_create(rx, ry, 0, ctypes.byref(h)) == nat.FI_E_INVALID_ARGUMENT
        assert "square radii" in nat.last_error()
    assert lib.fi_fftcc_create(40, 40, 0, ctypes.byref(h)) == nat.FI_E_UNSUPPORTED
    assert lib.fi_icgn_create(0, 5, 0.001, 10, 0, ctypes.byref(h)) == nat.FI_E_INVALID_ARGUMENT
    assert lib.fi_icgn_create(32, 32, 0.001, 10, 0, ctypes.byref(h)) == nat.FI_E_UNSUPPORTED
    assert lib.fi_fftcc_compute(None, None, 0) == nat.FI_E_INVALID_ARGUMENT
    assert lib.fi_frame_create(0, 0.001, 10, 0, ctypes.byref(h)) == nat.FI_E_INVALID_ARGUMENT


def test_python_api_maps_errors():
    from fasticgn_amd import FFTCC2D
    with pytest.raises(ValueError):
        FFTCC2D(7, 15)
    with pytest.raises(NotImplementedError):
        FFTCC2D(33, 33)


def test_raw_record_format():
    """formatRawBinary (opencorr_cli_cuda_mvp.cpp:2498-2536): x, y int32 then
    u, v, ux, uy, vx, vy, zncc float32, little endian, 36 bytes."""
    from fasticgn_amd import format_raw, poi_queue
    q = poi_queue([(17, 19), (22.0, 1017.0)])
    q["deformation"][0, [0, 1, 2, 6, 7, 8]] = [1.5, 0.25, -0.125, -2.5, 0.5, 0.75]
    q["result"][0, 2] = 0.99
    q["result"][1, 2] = -3
    raw = format_raw(q)
    assert len(raw) == 72
    rec = struct.unpack("<iifffffff", raw[:36])
    assert rec == (17, 19, 1.5, -2.5, 0.25, -0.125, 0.5, 0.75, pytest.approx(0.99))
    rec = struct.unpack("<iifffffff", raw[36:])
    assert rec[:2] == (22, 1017) and rec[-1] == -3
    assert format_raw(q[:0]) == b""
