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

consensus restates the neighbourhood consensus of the seeds
(launch_seed_consensus in fasticgn_amd/csrc/k_policy.hip): the same cells and
the same FP32 operations in the same order, so the device's seeds are
reproduced bit for bit; it is pinned by closed forms in
tests/test_seed_consensus.py (a linear field carried exactly over any
neighbourhood, the plain window mean of a symmetric one, a step excluded).

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


def seed_sample(cell: np.ndarray, ncx: int, mode: int) -> np.ndarray:
    """The POIs the network seeds under the consensus: every one (mode 0),
    those in cells (cx, cy) with cx + cy even (1), or cx and cy even (2)."""
    if mode == 0:
        return np.ones(len(cell), dtype=bool)
    cx, cy = cell % max(ncx, 1), cell // max(ncx, 1)
    pick = ((cx + cy) % 2 == 0) if mode == 1 else ((cx % 2 == 0) & (cy % 2 == 0))
    return (cell >= 0) & pick


def seed_cells(x: np.ndarray, y: np.ndarray, radius: float):
    """Cells of the consensus (build_seed_cells in fasticgn_amd/csrc/fi_runtime.cpp):
    side max(radius / 3, sqrt(area / m)) over the finite positions' bounding
    box in FP64, searched k = ceil(radius / side) cells to each side; returns
    (cell per position or -1, ncx, ncy, k, lists of positions per cell in
    position order)."""
    x = np.asarray(x, dtype=np.float32).astype(np.float64)
    y = np.asarray(y, dtype=np.float32).astype(np.float64)
    fin = np.isfinite(x) & np.isfinite(y)
    cell = np.full(len(x), -1, dtype=np.int64)
    m = int(fin.sum())
    if m == 0:
        return cell, 0, 0, 1, []
    x0, x1, y0, y1 = x[fin].min(), x[fin].max(), y[fin].min(), y[fin].max()
    r = float(np.float32(radius))
    side = max(r / 3.0, float(np.sqrt((x1 - x0 + 1.0) * (y1 - y0 + 1.0) / m)))
    k = max(1, int(np.ceil(r / side)))
    ncx = int(np.floor((x1 - x0) / side)) + 1
    ncy = int(np.floor((y1 - y0) / side)) + 1
    cx = np.minimum(np.floor((x[fin] - x0) / side).astype(np.int64), ncx - 1)
    cy = np.minimum(np.floor((y[fin] - y0) / side).astype(np.int64), ncy - 1)
    cell[fin] = cy * ncx + cx
    members = [[] for _ in range(ncx * ncy)]
    for i in np.nonzero(fin)[0]:
        members[cell[i]].append(int(i))
    return cell, ncx, ncy, k, [np.array(mb, dtype=np.int64) for mb in members]


def consensus(x: np.ndarray, y: np.ndarray, seeds: np.ndarray, ok: np.ndarray, radius: float, tau: float,
              sample_mode: int = 0, return_ok: bool = False):
    """Consensus seeds (n x 6: u, v, ux, uy, vx, vy) from the network's seeds:
    each POI with ok takes the mean of its ok neighbours within +-radius (x and
    y) whose predictions are finite, each neighbour's u, v carried to it with
    that neighbour's gradients,
    over the neighbours whose carried u, v agree with its own within tau; the
    mean gradients carry the mean u, v from the neighbours' centroid to it.
    FP32, sums in cell order (row by row around the POI's cell), then position
    order.  sample_mode (seed_sample): only the sampled POIs' predictions are
    used; a POI outside the sample (its own seeds and ok ignored) measures the
    tau test from the mean of the carried predictions within +-radius and is
    seeded (ok) when any neighbour then survives.  return_ok: also the ok of
    every POI after the consensus."""
    f = np.float32
    x = np.asarray(x, dtype=f)
    y = np.asarray(y, dtype=f)
    s = np.asarray(seeds, dtype=f)
    ok = np.asarray(ok, dtype=bool)
    out = s.copy()
    ok_out = ok.copy()
    cell, ncx, ncy, K, members = seed_cells(x, y, radius)
    sampled = seed_sample(cell, ncx, sample_mode)
    usable = sampled & ok & np.all(np.isfinite(s), axis=1)
    R, T = f(radius), f(tau)
    # sequential FP32 sums (cumsum is a running sum in the array's type)
    tot = lambda a: np.cumsum(a.astype(f), dtype=f)[-1] if len(a) else f(0)
    for i in range(len(x)):
        own = bool(sampled[i])
        if own and not ok[i]:
            continue
        if cell[i] < 0:
            if not own:
                ok_out[i] = False
            continue
        cx, cy = int(cell[i] % ncx), int(cell[i] // ncx)
        cand = [members[yy * ncx + xx] for yy in range(max(cy - K, 0), min(cy + K, ncy - 1) + 1)
                for xx in range(max(cx - K, 0), min(cx + K, ncx - 1) + 1)]
        j = np.concatenate(cand) if cand else np.zeros(0, dtype=np.int64)
        j = j[usable[j]]
        dx = (x[j] - x[i]).astype(f)
        dy = (y[j] - y[i]).astype(f)
        with np.errstate(invalid="ignore"):
            keep = (np.abs(dx) <= R) & (np.abs(dy) <= R)
            j, dx, dy = j[keep], dx[keep], dy[keep]
            sj = s[j]
            tu = (sj[:, 0] - ((sj[:, 2] * dx).astype(f) + (sj[:, 3] * dy).astype(f)).astype(f)).astype(f)
            tv = (sj[:, 1] - ((sj[:, 4] * dx).astype(f) + (sj[:, 5] * dy).astype(f)).astype(f)).astype(f)
            if own:
                au, av = s[i, 0], s[i, 1]
            else:
                if len(j) == 0:
                    ok_out[i] = False
                    continue
                inv1 = f(f(1.0) / f(len(j)))
                au, av = f(tot(tu) * inv1), f(tot(tv) * inv1)
            keep = (np.abs((tu - au).astype(f)) <= T) & (np.abs((tv - av).astype(f)) <= T)
        sj, dx, dy = sj[keep], dx[keep], dy[keep]
        if len(sj) == 0:
            if not own:
                ok_out[i] = False
            continue
        if not own:
            ok_out[i] = True
        cnt = f(len(sj))
        inv = f(f(1.0) / cnt)
        su, sv, sdx, sdy = tot(sj[:, 0]), tot(sj[:, 1]), tot(dx), tot(dy)
        g = [f(tot(sj[:, k]) * inv) for k in (2, 3, 4, 5)]
        mdx, mdy = f(sdx * inv), f(sdy * inv)
        out[i, 0] = f(f(su * inv) - f(f(g[0] * mdx) + f(g[1] * mdy)))
        out[i, 1] = f(f(sv * inv) - f(f(g[2] * mdx) + f(g[3] * mdy)))
        out[i, 2:] = g
    return (out, ok_out) if return_ok else out
