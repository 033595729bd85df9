"""CPU forward pass of the learned seeder (test infrastructure only).

Restates the device network of fasticgn_amd/csrc/fi_seeder.h in numpy: layer
l computes Z = X W + b with X and W rounded to BF16 (round to nearest even)
and FP32 accumulation, Y = max(Z, 0), and merges 2 x 2 patches of Y into the
next layer's rows (column block (dy * 2 + dx) * N + channel); the head is FP32
H W5 + b5.  The network has no reference counterpart; this restatement is
pinned by the tests against closed forms (linearity of the head, identity
layers, the patch permutation applied by hand) and is the checker of the
device kernels, not part of the product.
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
