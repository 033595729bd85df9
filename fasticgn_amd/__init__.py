"""MI355X (gfx950) implementation of the 2D-DIC hot path of FastICGN/OpenCorr:
FFTCC2D integer-pixel seeding, ICGN2D1 first-order IC-GN refinement over a
bicubic B-spline target, and the resident multi-frame driver.

The compute path is libfasticgn.so (hand-written HIP kernels behind the C ABI
in include/fasticgn.h).  Importing this package loads it and fails loudly when
it is not built; nothing here computes on the CPU.
"""
from ._native import POI_DTYPE, FrameStats, LogicError, DeviceError, device_count, load
from .opencorr import FFTCC2D, ICGN2D1, Image2D, poi_queue, grid_queue, format_raw
from .driver import FrameDriver, DEFAULT_POLICY, SKIP_FFTCC, POI_FALLBACK

load()

__all__ = [
    "POI_DTYPE", "FrameStats", "LogicError", "DeviceError", "device_count", "load",
    "FFTCC2D", "ICGN2D1", "Image2D", "poi_queue", "grid_queue", "format_raw",
    "FrameDriver", "DEFAULT_POLICY", "SKIP_FFTCC", "POI_FALLBACK",
]
