"""Learned seeder: Python side of the device network (include/fasticgn.h).

    LearnedSeeder      handle over fi_seeder_* (parameters, training, seeding)
    crop_rows          the network's layer-1 input for POIs of an image pair,
                       bit-identical to the fused inference kernel's A fragments
                       up to the order of the crop mean/variance sums
    make_training_set  block-mosaic speckle pairs, one affine map per 64x64
                       block (u, v ~ U[-8, 8] px, gradients ~ U[-0.05, 0.05]),
                       labelled by ICGN2D1 converged from the true map
    train              Adam over the set on the device kernels

Outputs are scaled: (u, v) in pixels relative to the POI's rounded initial
guess, gradients x GRAD_SCALE.  The trained parameter vector ships as
seeder_weights.npy beside this file (written by seeder_train.py).
"""
from __future__ import annotations

import ctypes
import math
import time
from pathlib import Path

import numpy as np

from . import _native as nat
from ._native import POI_DTYPE

CROP = 32
CROP_ORIGIN = 15
GRAD_SCALE = 100.0
N_OUT = 6
LAYERS = [(32, 16), (64, 32), (128, 64), (256, 128)]
WEIGHTS_PATH = Path(__file__).resolve().parent / "seeder_weights.npy"


def param_count() -> int:
    n = sum(k * m + m for k, m in LAYERS)
    return n + 128 * N_OUT + N_OUT


def unpack_params(p: np.ndarray):
    """[(W (K x N), b)] * 4 + [(W5 (128 x 6), b5)] views of the flat vector."""
    out, o = [], 0
    for k, m in LAYERS + [(128, N_OUT)]:
        w = p[o:o + k * m].reshape(k, m)
        o += k * m
        b = p[o:o + m]
        o += m
        out.append((w, b))
    assert o == len(p)
    return out


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 bit patterns, round to nearest even (finite values)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_value(bits: np.ndarray) -> np.ndarray:
    return (bits.astype(np.uint32) << 16).view(np.float32)


def crop_rows(ref: np.ndarray, tar: np.ndarray, pois: np.ndarray):
    """Layer-1 rows (n, 64, 32) as BF16 bits, the validity mask and the integer
    guess (n, 2).  Crops: 32x32 from (floor(x) - 15, floor(y) - 15) in the
    reference and shifted by (rint(u), rint(v)) in the target; each normalised
    to zero mean and unit RMS (FP32); row py*8 + px, column c*16 + dy*4 + dx
    holds crop c at (4py + dy, 4px + dx)."""
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    tar = np.ascontiguousarray(tar, dtype=np.float32)
    h, w = ref.shape
    n = len(pois)
    gu = np.rint(pois["deformation"][:, 0]).astype(np.float32)
    gv = np.rint(pois["deformation"][:, 6]).astype(np.float32)
    cx = np.floor(pois["x"]).astype(np.float32)
    cy = np.floor(pois["y"]).astype(np.float32)
    rx0, ry0 = cx - CROP_ORIGIN, cy - CROP_ORIGIN
    tx0, ty0 = rx0 + gu, ry0 + gv
    with np.errstate(invalid="ignore"):
        ok = ((rx0 >= 0) & (rx0 <= w - CROP) & (ry0 >= 0) & (ry0 <= h - CROP) &
              (tx0 >= 0) & (tx0 <= w - CROP) & (ty0 >= 0) & (ty0 <= h - CROP))
    x1 = np.zeros((n, 64, 32), dtype=np.uint16)
    idx = np.nonzero(ok)[0]
    if len(idx):
        ar = np.arange(CROP)
        crops = []
        for img, x0, y0 in ((ref, rx0, ry0), (tar, tx0, ty0)):
            yy = y0[idx].astype(np.int64)[:, None, None] + ar[None, :, None]
            xx = x0[idx].astype(np.int64)[:, None, None] + ar[None, None, :]
            c = img[yy, xx]  # (m, 32, 32)
            s = c.reshape(len(idx), -1).sum(axis=1, dtype=np.float32)
            sq = (c * c).reshape(len(idx), -1).sum(axis=1, dtype=np.float32)
            m = (s * np.float32(1.0 / 1024.0)).astype(np.float32)
            var = np.maximum(sq * np.float32(1.0 / 1024.0) - m * m, np.float32(0))
            inv = (np.float32(1.0) / np.sqrt(var + np.float32(1e-6))).astype(np.float32)
            crops.append(((c - m[:, None, None]) * inv[:, None, None]).astype(np.float32))
        # (m, c, py, dy, px, dx) -> rows (py, px), columns (c, dy, dx)
        st = np.stack(crops, axis=1).reshape(len(idx), 2, 8, 4, 8, 4)
        rows = st.transpose(0, 2, 4, 1, 3, 5).reshape(len(idx), 64, 32)
        x1[idx] = bf16_bits(rows)
    return x1, ok, np.stack([gu, gv], axis=1)


def targets_from(q: np.ndarray, guess: np.ndarray) -> np.ndarray:
    d = q["deformation"]
    return np.stack([d[:, 0] - guess[:, 0], d[:, 6] - guess[:, 1], GRAD_SCALE * d[:, 1], GRAD_SCALE * d[:, 2],
                     GRAD_SCALE * d[:, 7], GRAD_SCALE * d[:, 8]], axis=1).astype(np.float32)


class LearnedSeeder:
    """fi_seeder handle.  params: None = the shipped weights if present (else
    the seeded initialisation), or a flat FP32 vector."""

    def __init__(self, params: np.ndarray | None | str = None, device: int = 0):
        self._h = ctypes.c_void_p()
        nat.call("fi_seeder_create", int(device), ctypes.byref(self._h))
        if params is None and WEIGHTS_PATH.exists():
            params = np.load(WEIGHTS_PATH)
        if params is not None:
            self.set_params(np.asarray(params, dtype=np.float32))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            nat.call("fi_seeder_destroy", self._h)
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

    @property
    def handle(self):
        return self._h

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        nat.call("fi_seeder_set_params", self._h, nat.ptr(p), p.size)

    def get_params(self) -> np.ndarray:
        p = np.empty(param_count(), dtype=np.float32)
        nat.call("fi_seeder_get_params", self._h, nat.ptr(p), p.size)
        return p

    def set_training_data(self, x1: np.ndarray, targets: np.ndarray) -> None:
        x1 = np.ascontiguousarray(x1, dtype=np.uint16)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        assert x1.shape[1:] == (64, 32) and t.shape == (len(x1), N_OUT)
        nat.call("fi_seeder_set_training_data", self._h, nat.ptr(x1), nat.ptr(t), len(x1))

    def train_epoch(self, batch: int, lr: float, seed: int) -> float:
        loss = ctypes.c_float(0.0)
        nat.call("fi_seeder_train_epoch", self._h, int(batch), float(lr), ctypes.c_uint64(seed),
                 ctypes.byref(loss))
        return loss.value

    def gradient(self, x1: np.ndarray, targets: np.ndarray) -> tuple[np.ndarray, float]:
        """(parameter gradient, loss) of the training step on one batch (a
        multiple of 16 samples); the parameters are unchanged."""
        x1 = np.ascontiguousarray(x1, dtype=np.uint16)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        assert x1.shape[1:] == (64, 32) and t.shape == (len(x1), N_OUT)
        g = np.empty(param_count(), dtype=np.float32)
        loss = ctypes.c_float(0.0)
        nat.call("fi_seeder_gradient", self._h, nat.ptr(x1), nat.ptr(t), len(x1), nat.ptr(g), ctypes.byref(loss))
        return g, loss.value

    def forward(self, x1: np.ndarray) -> np.ndarray:
        x1 = np.ascontiguousarray(x1, dtype=np.uint16)
        out = np.empty((len(x1), N_OUT), dtype=np.float32)
        nat.call("fi_seeder_forward", self._h, nat.ptr(x1), len(x1), nat.ptr(out))
        return out

    def seed(self, ref: np.ndarray, tar: np.ndarray, pois: np.ndarray,
             consensus: tuple[float, float] | None = None, passes: int = 1, sample: str = "all") -> np.ndarray:
        """Seeds POIs in place (fused inference kernel); returns the ok mask.
        consensus = (radius, tau): then the neighbourhood consensus of the
        predictions (fi_seeder_seed_consensus), and passes - 1 further passes
        each followed by it, as the frame driver seeds with its subset radius;
        sample: the POIs the network seeds (FrameDriver.set_seed_consensus)."""
        ref = np.ascontiguousarray(ref, dtype=np.float32)
        tar = np.ascontiguousarray(tar, dtype=np.float32)
        assert ref.shape == tar.shape and pois.dtype == POI_DTYPE and pois.flags.c_contiguous
        ok = np.zeros(len(pois), dtype=np.uint32)
        h, w = ref.shape
        if consensus is None:
            if passes != 1 or sample != "all":
                raise ValueError("further seeder passes and samples go with the consensus (the frame driver's seeding)")
            nat.call("fi_seeder_seed", self._h, nat.ptr(ref), nat.ptr(tar), w, h, nat.ptr(pois), len(pois),
                     nat.ptr(ok))
        else:
            radius, tau = consensus
            nat.call("fi_seeder_seed_consensus", self._h, nat.ptr(ref), nat.ptr(tar), w, h, nat.ptr(pois),
                     len(pois), nat.ptr(ok), float(radius), float(tau), int(passes), nat.SEED_SAMPLES[sample])
        return ok.astype(bool)


# ---------------------------------------------------------------------------
# training data
# ---------------------------------------------------------------------------
BLOCK = 64


def block_mosaic(blocks_per_side: int, rng: np.random.Generator, max_shift: float = 8.0,
                 max_grad: float = 0.05, dot_density: float = 6000.0 / 512 ** 2):
    """Reference speckle of (blocks_per_side * 64)^2 pixels and a target whose
    64x64 block b follows its own affine map about the block centre c_b:
    x = c_b + A_b (X - c_b) + t_b.  Returns ref, tar (u8), POIs at the block
    centres and the true (u, v, ux, uy, vx, vy) per block."""
    from .synth import cubic_resample, dot_field, quantize, render_dots
    from .opencorr import poi_queue
    side = blocks_per_side * BLOCK
    cx, cy, s = dot_field(side, side, int(dot_density * side * side), rng)
    ref = quantize(render_dots(cx, cy, s, side, side, 40.0, 140.0), rng, 1.5, np.uint8, 255.0)
    nb = blocks_per_side * blocks_per_side
    t = rng.uniform(-max_shift, max_shift, (nb, 2))
    g = rng.uniform(-max_grad, max_grad, (nb, 4))  # ux, uy, vx, vy
    A = np.empty((nb, 2, 2))
    A[:, 0, 0] = 1 + g[:, 0]
    A[:, 0, 1] = g[:, 1]
    A[:, 1, 0] = g[:, 2]
    A[:, 1, 1] = 1 + g[:, 3]
    Ai = np.linalg.inv(A)
    by, bx = np.divmod(np.arange(nb), blocks_per_side)
    ccx = bx * BLOCK + BLOCK // 2
    ccy = by * BLOCK + BLOCK // 2
    ys, xs = np.mgrid[0:side, 0:side]
    b = (ys // BLOCK) * blocks_per_side + (xs // BLOCK)
    dx = xs - ccx[b] - t[b, 0]
    dy = ys - ccy[b] - t[b, 1]
    X = ccx[b] + Ai[b, 0, 0] * dx + Ai[b, 0, 1] * dy
    Y = ccy[b] + Ai[b, 1, 0] * dx + Ai[b, 1, 1] * dy
    tar = quantize(cubic_resample(ref, X, Y), rng, 1.5, np.uint8, 255.0)
    pois = poi_queue(np.stack([ccx, ccy], axis=1).astype(np.float32))
    truth = np.concatenate([t, g], axis=1)  # u, v, ux, uy, vx, vy
    truth = truth[:, [0, 1, 2, 3, 4, 5]]
    return ref, tar, pois, truth


def make_training_set(n_samples: int, seed: int, blocks_per_side: int = 32, device: int = 0,
                      log=print):
    """(x1 (n, 64, 32) BF16 bits, targets (n, 6), stats).  Labels: ICGN2D1
    (r = 15, conv 1e-4, max 20) converged from the true map; samples whose
    IC-GN result is not valid with ZNCC >= 0.9 or strays more than 0.25 px from
    the truth are dropped."""
    from .opencorr import ICGN2D1, Image2D
    rng = np.random.default_rng(seed)
    xs, ts, kept, total = [], [], 0, 0
    t0 = time.perf_counter()
    while kept < n_samples:
        ref, tar, pois, truth = block_mosaic(blocks_per_side, rng)
        q = pois.copy()
        q["deformation"][:, 0] = truth[:, 0]
        q["deformation"][:, 6] = truth[:, 1]
        q["deformation"][:, 1] = truth[:, 2]
        q["deformation"][:, 2] = truth[:, 3]
        q["deformation"][:, 7] = truth[:, 4]
        q["deformation"][:, 8] = truth[:, 5]
        with ICGN2D1(15, 15, 1e-4, 20, device) as ic:
            ic.setImages(Image2D(ref), Image2D(tar))
            ic.prepare()
            ic.compute(q)
        z = q["result"][:, 2]
        good = np.isfinite(z) & (z >= 0.9)
        good &= np.abs(q["deformation"][:, 0] - truth[:, 0]) < 0.25
        good &= np.abs(q["deformation"][:, 6] - truth[:, 1]) < 0.25
        crops_q = pois.copy()  # guess 0: the target crop sits at the POI
        x1, ok, guess = crop_rows(ref.astype(np.float32), tar.astype(np.float32), crops_q)
        good &= ok
        xs.append(x1[good])
        ts.append(targets_from(q[good], guess[good]))
        kept += int(good.sum())
        total += len(q)
        log(f"[seeder] data {kept}/{n_samples} ({total} generated, {time.perf_counter() - t0:.1f} s)")
    x1 = np.concatenate(xs)[:n_samples]
    t = np.concatenate(ts)[:n_samples]
    return x1, t, {"generated": total, "kept": kept, "seconds": time.perf_counter() - t0}


def train(seeder: LearnedSeeder, x1: np.ndarray, targets: np.ndarray, epochs: int = 20, batch: int = 256,
          lr: float = 2e-3, seed: int = 1, log=print) -> list:
    seeder.set_training_data(x1, targets)
    losses = []
    for e in range(epochs):
        rate = lr * 0.5 * (1 + math.cos(math.pi * e / epochs))  # cosine decay
        t0 = time.perf_counter()
        loss = seeder.train_epoch(batch, rate, seed + e)
        losses.append(loss)
        log(f"[seeder] epoch {e + 1}/{epochs} lr {rate:.2e} loss {loss:.5f} ({time.perf_counter() - t0:.2f} s)")
    return losses
