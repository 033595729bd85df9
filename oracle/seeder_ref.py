"""CPU forward and backward passes of the learned seeder (test
infrastructure only).

Restates the device network of fasticgn_amd/csrc/fi_seeder.h in numpy: layer
l computes Z = X W + b with X and W rounded to BF16 (round to nearest even)
and FP32 accumulation, Y = max(Z, 0), and merges 2 x 2 patches of Y into the
next layer's rows (column block (dy * 2 + dx) * N + channel); the head is FP32
H W5 + b5.  The network has no reference counterpart; this restatement is
pinned by the tests against closed forms (linearity of the head, identity
layers, the patch permutation applied by hand) and is the checker of the
device kernels, not part of the product.

loss_and_grad restates one training step's loss and parameter gradient with
the device's rounding points (BF16 inputs of every GEMM, FP32 elsewhere); with
bf16=False it is the plain float64 backward pass, pinned by finite
differences in tests/test_seeder.py.
"""
from __future__ import annotations

import numpy as np

LAYERS = [(32, 16), (64, 32), (128, 64), (256, 128)]
GRIDS = [8, 4, 2, 1]
N_OUT = 6


def bf16_round(x: np.ndarray) -> np.ndarray:
    """FP32 values rounded to BF16 (nearest even), returned as FP32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def bits_to_f32(bits: np.ndarray) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def unpack(params: np.ndarray):
    out, o = [], 0
    for k, n in LAYERS + [(128, N_OUT)]:
        w = params[o:o + k * n].reshape(k, n)
        o += k * n
        out.append((w, params[o:o + n]))
        o += n
    assert o == params.size
    return out


def merge_patches(y: np.ndarray, samples: int, g: int) -> np.ndarray:
    """(samples * g * g, N) rows on a g x g grid -> (samples * (g/2)^2, 4N)."""
    n = y.shape[1]
    t = y.reshape(samples, g // 2, 2, g // 2, 2, n).transpose(0, 1, 3, 2, 4, 5)
    return t.reshape(samples * (g // 2) ** 2, 4 * n)


def forward(params: np.ndarray, x1_bits: np.ndarray) -> np.ndarray:
    """Scaled outputs (n, 6) for layer-1 rows (n, 64, 32) of BF16 bits."""
    params = np.asarray(params, dtype=np.float32)
    layers = unpack(params)
    n = len(x1_bits)
    x = bits_to_f32(x1_bits).reshape(n * 64, 32)
    h = None
    for l, (w, b) in enumerate(layers[:4]):
        z = x.astype(np.float32) @ bf16_round(w).astype(np.float32) + b.astype(np.float32)
        y = np.maximum(z, np.float32(0)).astype(np.float32)
        if l < 3:
            x = bf16_round(merge_patches(y, n, GRIDS[l]))
        else:
            h = y
    w5, b5 = layers[4]
    return (h @ w5.astype(np.float32) + b5).astype(np.float32)


def unmerge_patches(d: np.ndarray, samples: int, g: int) -> np.ndarray:
    """Inverse of merge_patches: (samples * (g/2)^2, 4N) -> (samples * g * g, N)."""
    n = d.shape[1] // 4
    t = d.reshape(samples, g // 2, g // 2, 2, 2, n).transpose(0, 1, 3, 2, 4, 5)
    return t.reshape(samples * g * g, n)


def loss_and_grad(params: np.ndarray, x1_bits: np.ndarray, targets: np.ndarray, bf16: bool = True):
    """(loss, gradient) of the mean squared error over the n x 6 scaled
    outputs.  bf16: round every GEMM input (layer inputs, weights, the ReLU'
    gradients that feed the weight and input GEMMs) as the training kernels
    do, in FP32; otherwise float64 throughout, no rounding."""
    dt = np.float32 if bf16 else np.float64
    R = bf16_round if bf16 else (lambda a: a)
    P = np.asarray(params, dtype=dt)
    layers = unpack(P)
    n = len(x1_bits)
    t = np.asarray(targets, dtype=dt)
    xs, zbs = [], []
    x = bits_to_f32(x1_bits).reshape(n * 64, 32).astype(dt)
    for l, (w, b) in enumerate(layers[:4]):
        xs.append(x)
        zb = (x @ R(w).astype(dt) + b).astype(dt)
        zbs.append(zb)
        y = np.maximum(zb, dt(0))
        if l < 3:
            x = R(merge_patches(y, n, GRIDS[l]).astype(np.float32)).astype(dt) if bf16 else merge_patches(y, n, GRIDS[l])
    h = np.maximum(zbs[3], dt(0))
    w5, b5 = layers[4]
    out = h @ w5 + b5
    d = out - t
    loss = float((d.astype(np.float64) ** 2).mean())
    grads = []
    dout = (dt(2) * d / dt(n * N_OUT)).astype(dt)
    g5w, g5b = h.T @ dout, dout.sum(axis=0)
    dnext = dout @ w5.T
    per_layer = [None] * 4
    for l in range(3, -1, -1):
        w, _ = layers[l]
        dy = dnext if l == 3 else unmerge_patches(dnext, n, GRIDS[l])
        dz = np.where(zbs[l] > 0, dy, dt(0)).astype(dt)
        dzr = R(dz.astype(np.float32)).astype(dt) if bf16 else dz
        gw = xs[l].T @ dzr
        gb = dz.sum(axis=0)
        per_layer[l] = (gw, gb)
        if l > 0:
            dnext = dzr @ R(w).astype(dt).T
    for gw, gb in per_layer:
        grads += [gw.ravel(), gb]
    grads += [g5w.ravel(), g5b]
    return loss, np.concatenate(grads).astype(dt)
