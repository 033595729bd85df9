"""OpenCorr-shaped Python API over the gfx950 C ABI.

Names, argument meaning and error behaviour follow the reference's C++ classes
so that code written against them reads the same:

    FFTCC2D(rx, ry, ...)                   src/oc_fftcc.h:56-100, src/oc_fftcc.cpp:154-557
    ICGN2D1(rx, ry, conv, stop, ...)       src/oc_icgn.h:45-77,   src/oc_icgn.cpp:95-397
    POI2D queue                            src/oc_poi.h:102-136 (numpy structured array)

Differences that the device path imposes are explicit: radius limits raise
NotImplementedError (FI_E_UNSUPPORTED), and ICGN2D1.compute requires
prepare() after the images change (the reference would read stale tables).
A POI queue is a numpy array of POI_DTYPE; compute() updates it in place.
"""
from __future__ import annotations

import ctypes
from typing import Iterable, Sequence

import numpy as np

from . import _native as nat
from ._native import POI_DTYPE


def poi_queue(points: Iterable[Sequence[float]] | np.ndarray) -> np.ndarray:
    """POI2D(x, y) for each point, every other field zero (POI2D::clear)."""
    pts = np.asarray(points, dtype=np.float32).reshape(-1, 2)
    q = np.zeros(len(pts), dtype=POI_DTYPE)
    q["x"] = pts[:, 0]
    q["y"] = pts[:, 1]
    return q


def grid_queue(width: int, height: int, radius: int, grid: int, margin: int | None = None) -> np.ndarray:
    """Integer POI grid of buildPoiQueue (opencorr_cli_cuda_mvp.cpp:2439-2472):
    x, y from margin (default radius + 2) while < size - margin, step grid."""
    m = radius + 2 if margin is None else margin
    xs = np.arange(m, width - m, grid, dtype=np.float32)
    ys = np.arange(m, height - m, grid, dtype=np.float32)
    yy, xx = np.meshgrid(ys, xs, indexing="ij")
    return poi_queue(np.stack([xx.ravel(), yy.ravel()], axis=1))


def _check_queue(queue: np.ndarray) -> np.ndarray:
    if not isinstance(queue, np.ndarray) or queue.dtype != POI_DTYPE:
        raise TypeError("POI queue must be a numpy array of POI_DTYPE")
    if not queue.flags.c_contiguous:
        raise ValueError("POI queue must be C-contiguous")
    return queue


class Image2D:
    """A grey image: row-major uint8, uint16 or float32 pixels, and the input
    mode that converts them (loadImage2DMat, src/oc_image.cpp:125-217).

    depth12: 'native' (the stored type's depth), 'right' (12-bit, values must
    be <= 4095), 'left' (12-bit, divided by 16), 'auto12' (12-bit,
    left-aligned when the maximum exceeds 4095), '16' (16-bit) or 'auto'
    (16-bit data up to 4095 is 12-bit right-aligned, else 16-bit; 8-bit data
    8-bit).  normalize: 'none',
    'to8', 'to16' or 'unit' (0..255, 0..65535 or 0..1 from the logical maximum
    of the effective depth)."""

    _DEPTH = {"native": nat.FI_DEPTH_NATIVE, "right": nat.FI_DEPTH12_RIGHT, "left": nat.FI_DEPTH12_LEFT,
              "auto12": nat.FI_DEPTH12_AUTO, "auto": nat.FI_DEPTH_AUTO, "16": nat.FI_DEPTH16}
    _NORMALIZE = {"none": nat.FI_NORMALIZE_NONE, "to8": nat.FI_NORMALIZE_TO8, "to16": nat.FI_NORMALIZE_TO16,
                  "unit": nat.FI_NORMALIZE_UNIT}

    def __init__(self, pixels: np.ndarray, depth12: str = "native", normalize: str = "none"):
        a = np.ascontiguousarray(pixels)
        if a.ndim != 2:
            raise ValueError("Image2D needs a 2-D array")
        if a.dtype == np.uint8:
            self.pixel = nat.FI_PIXEL_U8
        elif a.dtype == np.uint16:
            self.pixel = nat.FI_PIXEL_U16
        elif a.dtype == np.float32:
            self.pixel = nat.FI_PIXEL_F32
        else:
            raise TypeError(f"unsupported pixel type {a.dtype}")
        if depth12 not in self._DEPTH:
            raise ValueError(f"depth12 must be one of {sorted(self._DEPTH)}")
        if normalize not in self._NORMALIZE:
            raise ValueError(f"normalize must be one of {sorted(self._NORMALIZE)}")
        self.data = a
        self.mode = self._DEPTH[depth12] | self._NORMALIZE[normalize]
        self.height, self.width = a.shape
        self.file_path = ""

    @property
    def depth12(self) -> int:
        return self.mode

    def args(self):
        return (nat.ptr(self.data), self.width, self.height, self.data.strides[0], self.pixel, self.mode)

    def info(self) -> dict:
        """The effective depth, alignment and scale of this image's mode
        (ImageLoadInfo), as an upload resolves them (fi_image_resolve)."""
        out = nat.IngestInfo()
        nat.call("fi_image_resolve", *self.args(), ctypes.byref(out))
        return out.as_dict()

    def as_float(self) -> np.ndarray:
        """The pixels as the device converts them: float(p) * float(scale)."""
        scale = np.float32(self.info()["scale"])
        return (self.data.astype(np.float32) * scale).astype(np.float32)

    @classmethod
    def load(cls, path, depth: str = "compat8", alignment: str = "auto", normalize: str = "none") -> "Image2D":
        """Image2D(file_path) / loadWithOptions (src/oc_image.cpp:229-285) for
        PGM files (P2 / P5, 8- or 16-bit).  depth 'compat8' (the reference's
        default: 8-bit, 16-bit data reduced to its high byte), '12' (with
        alignment 'right', 'left' or 'auto'), '16' or 'auto'."""
        px = read_pgm(path)
        if depth == "compat8":
            if px.dtype == np.uint16:
                px = (px >> 8).astype(np.uint8)
            d = "native"
        elif depth == "12":
            d = {"right": "right", "left": "left", "auto": "auto12"}[alignment]
        elif depth == "16":
            d = "16"
        elif depth == "auto":
            d = "auto" if px.dtype == np.uint16 else "native"
        else:
            raise ValueError("depth must be 'compat8', '12', '16' or 'auto'")
        img = cls(px, d, normalize)
        img.file_path = str(path)
        return img


def read_pgm(path) -> np.ndarray:
    """Decodes a PGM file (P5 binary or P2 ASCII; maxval up to 255 gives
    uint8, up to 65535 big-endian uint16)."""
    blob = open(path, "rb").read()
    fields, pos = [], 0
    while len(fields) < 4:
        while pos < len(blob) and blob[pos:pos + 1].isspace():
            pos += 1
        if pos < len(blob) and blob[pos:pos + 1] == b"#":
            while pos < len(blob) and blob[pos:pos + 1] not in (b"\n", b"\r"):
                pos += 1
            continue
        start = pos
        while pos < len(blob) and not blob[pos:pos + 1].isspace():
            pos += 1
        if start == pos:
            raise ValueError(f"Fail to load file: {path}")
        fields.append(blob[start:pos])
    magic, w, h, maxval = fields[0], int(fields[1]), int(fields[2]), int(fields[3])
    if magic not in (b"P5", b"P2") or w < 1 or h < 1 or not 0 < maxval < 65536:
        raise ValueError(f"Fail to load file: {path}")
    dtype = np.uint8 if maxval < 256 else np.dtype(">u2")
    if magic == b"P5":
        n = w * h * np.dtype(dtype).itemsize
        data = np.frombuffer(blob, dtype=dtype, count=w * h, offset=pos + 1) if len(blob) >= pos + 1 + n else None
        if data is None:
            raise ValueError(f"Fail to load file: {path}")
    else:
        data = np.array(blob[pos:].split()[: w * h], dtype=np.int64)
        if data.size != w * h:
            raise ValueError(f"Fail to load file: {path}")
    return data.reshape(h, w).astype(np.uint8 if maxval < 256 else np.uint16)


def write_pgm(path, px: np.ndarray) -> None:
    """A binary PGM (P5) of uint8 or uint16 pixels."""
    px = np.ascontiguousarray(px)
    maxval = 255 if px.dtype == np.uint8 else 65535
    with open(path, "wb") as fh:
        fh.write(f"P5\n{px.shape[1]} {px.shape[0]}\n{maxval}\n".encode("ascii"))
        fh.write(px.astype(">u2").tobytes() if px.dtype == np.uint16 else px.tobytes())


class _Handle:
    _destroy = ""

    def __init__(self):
        self._h = ctypes.c_void_p()

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            nat.call(self._destroy, self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class FFTCC2D(_Handle):
    """opencorr::FFTCC2D on the device (FFTCC2DCuda's device argument)."""

    _destroy = "fi_fftcc_destroy"

    def __init__(self, subset_radius_x: int, subset_radius_y: int, device: int = 0):
        super().__init__()
        nat.call("fi_fftcc_create", int(subset_radius_x), int(subset_radius_y), int(device),
                 ctypes.byref(self._h))
        self.subset_radius_x = int(subset_radius_x)
        self.subset_radius_y = int(subset_radius_y)
        self.ref_img: Image2D | None = None
        self.tar_img: Image2D | None = None

    def setImages(self, ref: Image2D, tar: Image2D) -> None:
        self.setRefImage(ref)
        self.setTarImage(tar)

    def setRefImage(self, ref: Image2D) -> None:
        nat.call("fi_fftcc_set_ref", self._h, *ref.args())
        self.ref_img = ref

    def setTarImage(self, tar: Image2D) -> None:
        nat.call("fi_fftcc_set_tar", self._h, *tar.args())
        self.tar_img = tar

    def hasPreparedReference(self, ref: Image2D) -> bool:
        out = ctypes.c_int(0)
        nat.call("fi_fftcc_has_prepared_ref", self._h, ctypes.byref(out))
        return bool(out.value) and ref is self.ref_img

    def prepare(self) -> None:
        self.prepareRef()
        self.prepareTar()

    def prepareRef(self, queue: np.ndarray | None = None) -> None:
        if queue is None:
            nat.call("fi_fftcc_prepare_ref_clear", self._h)
            return
        q = _check_queue(queue)
        nat.call("fi_fftcc_prepare_ref", self._h, nat.ptr(q), len(q))

    def prepareTar(self) -> None:
        pass  # the target is read directly (src/oc_fftcc.cpp:315)

    def compute(self, queue: np.ndarray, cache_indices: Sequence[int] | None = None) -> None:
        q = _check_queue(queue)
        if cache_indices is None:
            nat.call("fi_fftcc_compute", self._h, nat.ptr(q), len(q))
            return
        idx = np.ascontiguousarray(cache_indices, dtype=np.uint64)
        if len(idx) != len(q):
            raise ValueError("FFTCC2D::compute requires one cache index per POI")
        nat.call("fi_fftcc_compute_indexed", self._h, nat.ptr(q), len(q), nat.ptr(idx))


class ICGN2D1(_Handle):
    """opencorr::ICGN2D1 on the device."""

    _destroy = "fi_icgn_destroy"

    def __init__(self, subset_radius_x: int, subset_radius_y: int, conv_criterion: float,
                 stop_condition: float, device: int = 0):
        super().__init__()
        nat.call("fi_icgn_create", int(subset_radius_x), int(subset_radius_y),
                 float(conv_criterion), float(stop_condition), int(device), ctypes.byref(self._h))
        self.subset_radius_x = int(subset_radius_x)
        self.subset_radius_y = int(subset_radius_y)
        self.conv_criterion = float(conv_criterion)
        self.stop_condition = float(stop_condition)
        self.self_adaptive = False

    def setIteration(self, conv_criterion: float, stop_condition: float) -> None:
        nat.call("fi_icgn_set_iteration", self._h, float(conv_criterion), float(stop_condition))
        self.conv_criterion, self.stop_condition = float(conv_criterion), float(stop_condition)

    def setSelfAdaptive(self, on: bool) -> None:
        nat.call("fi_icgn_set_self_adaptive", self._h, int(bool(on)))
        self.self_adaptive = bool(on)

    def setShape(self, waves: int, rows_per_band: int) -> None:
        """Lane-to-POI shape of the IC-GN kernel: waves per POI and rows per
        band (0, 0 = default)."""
        nat.call("fi_icgn_set_shape", self._h, int(waves), int(rows_per_band))

    def setSampling(self, mode: str) -> None:
        """'auto' (raw neighbourhood staged in LDS, table fallback), 'table'
        (every sample from the B-spline table) or 'group' (four neighbouring
        POIs share an LDS tile of x-polynomial coefficients; one-wave lane
        shapes, others run as 'auto')."""
        modes = nat.SAMPLING_MODES
        if mode not in modes:
            raise ValueError(f"sampling must be one of {sorted(modes)}")
        nat.call("fi_icgn_set_sampling", self._h, modes[mode])

    def setImages(self, ref: Image2D, tar: Image2D) -> None:
        self.setRefImage(ref)
        self.setTarImage(tar)

    def setRefImage(self, ref: Image2D) -> None:
        nat.call("fi_icgn_set_ref", self._h, *ref.args())

    def setTarImage(self, tar: Image2D) -> None:
        nat.call("fi_icgn_set_tar", self._h, *tar.args())

    def prepareRef(self) -> None:
        nat.call("fi_icgn_prepare_ref", self._h)

    def prepareTar(self) -> None:
        nat.call("fi_icgn_prepare_tar", self._h)

    def prepare(self) -> None:
        nat.call("fi_icgn_prepare", self._h)

    def compute(self, queue: np.ndarray, center_offsets: np.ndarray | None = None) -> None:
        """compute(queue), or with center_offsets (n x 2) compute(queue,
        center_offset_queue) (src/oc_icgn.cpp:399-625): POI i is matched with
        its warp centred at (x + ox, y + oy)."""
        q = _check_queue(queue)
        if center_offsets is None:
            nat.call("fi_icgn_compute", self._h, nat.ptr(q), len(q))
            return
        off = np.ascontiguousarray(center_offsets, dtype=np.float32)
        if off.shape != (len(q), 2):
            raise ValueError("center_offsets must be one (x, y) pair per POI")
        nat.call("fi_icgn_compute_offset", self._h, nat.ptr(q), len(q), nat.ptr(off))

    def export(self, what: str, shape=None, n_records: int = 0) -> np.ndarray:
        """A device-side array of the engine (fi_icgn_export): 'ref', 'gx',
        'gy', 'tar' (h x w f32), 'bspline' (h x w x 16 f32, coef[k][l] at
        k * 4 + l) with shape = (h, w), or 'prep' (the last compute's
        n_records reference records, PREP_DTYPE)."""
        kinds = {"ref": nat.FI_EXPORT_REF_F32, "gx": nat.FI_EXPORT_GRAD_X, "gy": nat.FI_EXPORT_GRAD_Y,
                 "tar": nat.FI_EXPORT_TAR_F32, "bspline": nat.FI_EXPORT_BSPLINE, "prep": nat.FI_EXPORT_PREP}
        if what not in kinds:
            raise ValueError(f"export must be one of {sorted(kinds)}")
        if what == "prep":
            out = np.zeros(n_records, dtype=nat.PREP_DTYPE)
        else:
            h, w = shape
            out = np.zeros((h, w, 16) if what == "bspline" else (h, w), dtype=np.float32)
        nat.call("fi_icgn_export", self._h, kinds[what], nat.ptr(out), out.nbytes)
        return out


def format_raw(queue: np.ndarray) -> bytes:
    """36-byte records x,y (int32) u,v,ux,uy,vx,vy,zncc (float32), little endian
    (formatRawBinary, opencorr_cli_cuda_mvp.cpp:2498-2536)."""
    q = _check_queue(queue)
    out = np.empty(len(q) * 36, dtype=np.uint8)
    written = ctypes.c_size_t(0)
    nat.call("fi_format_raw", nat.ptr(q), len(q), nat.ptr(out), out.nbytes, ctypes.byref(written))
    return out[: written.value].tobytes()
