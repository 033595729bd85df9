"""fi_icgn_compute_noalloc against fi_icgn_compute, through the C ABI: the
same POIs computed from host memory by the existing entry point and in place
in device memory (the test's own hipMalloc) by the entry point that allocates nothing,
on two calls in a row, the second with fewer POIs; results byte-identical
(same kernels, the records computed into the library's static array instead
of the handle's buffer).  Sizes beyond the array are rejected."""
import ctypes

import numpy as np
import pytest

import fasticgn_amd as fa
from fasticgn_amd import _native as nat
from fasticgn_amd import synth

pytestmark = pytest.mark.gpu


def hip():
    """The HIP runtime the library runs on, for the test's own device buffer."""
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return lib


def test_noalloc_matches_compute_twice():
    h = hip()
    sc = synth.config0()
    nbytes = len(sc.pois) * fa.POI_DTYPE.itemsize
    dev = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(dev), nbytes) == 0
    try:
        run_two_calls(h, dev, sc)
    finally:
        h.hipFree(dev)


def run_two_calls(h, dev, sc):
    with fa.ICGN2D1(15, 15, 0.001, 10) as ic:
        ic.setImages(fa.Image2D(sc.ref), fa.Image2D(sc.tar))
        ic.prepare()
        for n in (len(sc.pois), 700):  # the second call with a smaller input
            q = sc.pois[:n].copy()
            u, v, *_ = sc.truth(q["x"], q["y"])
            q["deformation"][:, 0] = np.round(u)
            q["deformation"][:, 6] = np.round(v)
            want = q.copy()
            ic.compute(want)
            assert h.hipMemcpy(dev, nat.ptr(q), q.nbytes, 1) == 0  # host to device
            nat.call("fi_icgn_compute_noalloc", ic._h, dev, n)
            got = np.zeros(n, dtype=fa.POI_DTYPE)
            assert h.hipMemcpy(nat.ptr(got), dev, got.nbytes, 2) == 0  # device to host
            print("no-alloc compute", n, "POIs, valid", int((got["result"][:, 2] > 0).sum()), flush=True)
            assert got.tobytes() == want.tobytes(), n
            assert (got["result"][:, 2] > 0.9).mean() > 0.9
        with pytest.raises(ValueError, match="at most 65536"):
            nat.call("fi_icgn_compute_noalloc", ic._h, dev, 65537)
        ic.setSelfAdaptive(True)
        with pytest.raises(NotImplementedError):
            nat.call("fi_icgn_compute_noalloc", ic._h, dev, 10)
