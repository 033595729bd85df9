"""Train the learned seeder on the device and write its parameter vector.

    python -m fasticgn_amd.seeder_train [--samples 200000] [--epochs 20]
        [--out fasticgn_amd/seeder_weights.npy]

Data: block-mosaic speckle pairs (one affine map per 64x64 block, u, v ~
U[-8, 8] px, gradients ~ U[-0.05, 0.05]), generated by a pool of CPU workers
and labelled by ICGN2D1 on the device (as seeder.make_training_set); the
target crop sits at the rounded truth moved by U[-1, 1] px for a fraction
--near of the samples (default 0.9: the regime of the second seeding pass and
of every POI a neighbourhood consensus has already pulled close) and at guess
0 for the rest.  5 % of the samples are held out and the prediction error on
them is reported for both groups.  Training
runs entirely in the BF16-MFMA kernels (fi_seeder_train_epoch).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np


def _mosaic(args):
    seed, blocks = args
    from fasticgn_amd.seeder import block_mosaic
    return block_mosaic(blocks, np.random.default_rng(seed))


def generate(n_samples: int, seed: int, workers: int, blocks: int = 32, near_fraction: float = 0.5,
             log=print):
    """Like seeder.make_training_set, with the mosaics rendered in parallel."""
    from fasticgn_amd.opencorr import ICGN2D1, Image2D
    from fasticgn_amd.seeder import crop_rows, targets_from
    per = blocks * blocks
    n_mosaics = int(np.ceil(n_samples / (0.9 * per))) + 2
    xs, ts, nears, kept, total = [], [], [], 0, 0
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=workers) as pool, ICGN2D1(15, 15, 1e-4, 20) as ic:
        for ref, tar, pois, truth in pool.map(_mosaic, [(seed * 100003 + m, blocks) for m in range(n_mosaics)]):
            q = pois.copy()
            for k, col in enumerate((0, 6, 1, 2, 7, 8)):
                q["deformation"][:, col] = truth[:, k]
            ic.setImages(Image2D(ref), Image2D(tar))
            ic.prepare()
            ic.compute(q)
            z = q["result"][:, 2]
            good = np.isfinite(z) & (z >= 0.9)
            good &= np.abs(q["deformation"][:, 0] - truth[:, 0]) < 0.25
            good &= np.abs(q["deformation"][:, 6] - truth[:, 1]) < 0.25
            # crop guesses: half at 0 (the full +-8 px range), half at the
            # rounded truth moved by U[-1, 1] px (the residual range of a
            # second pass or of a good initial guess)
            c = pois.copy()
            rng = np.random.default_rng(seed * 7919 + total)
            near = rng.random(len(c)) < near_fraction
            c["deformation"][:, 0] = np.where(near, np.rint(truth[:, 0] + rng.uniform(-1, 1, len(c))), 0)
            c["deformation"][:, 6] = np.where(near, np.rint(truth[:, 1] + rng.uniform(-1, 1, len(c))), 0)
            x1, ok, guess = crop_rows(ref.astype(np.float32), tar.astype(np.float32), c)
            good &= ok
            xs.append(x1[good])
            ts.append(targets_from(q[good], guess[good]))
            nears.append(near[good])
            kept += int(good.sum())
            total += len(q)
            log(f"[seeder] data {kept}/{n_samples} kept of {total} ({time.perf_counter() - t0:.1f} s)")
            if kept >= n_samples:
                break
    return (np.concatenate(xs)[:n_samples], np.concatenate(ts)[:n_samples],
            np.concatenate(nears)[:n_samples], total)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=200000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--workers", type=int, default=min(12, os.cpu_count() or 1))
    ap.add_argument("--out", default=str(Path(__file__).resolve().parent / "seeder_weights.npy"))
    ap.add_argument("--report", default="")
    ap.add_argument("--near", type=float, default=0.9, help="fraction of crops near the truth")
    args = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    from fasticgn_amd import seeder as sd
    x1, t, near, generated = generate(args.samples, args.seed, args.workers, near_fraction=args.near, log=log)
    n_hold = max(1, len(x1) // 20)
    rng = np.random.default_rng(args.seed)
    perm = rng.permutation(len(x1))
    hold, train_idx = perm[:n_hold], perm[n_hold:]
    t_start = time.perf_counter()
    with sd.LearnedSeeder(params=_init(sd)) as s:
        losses = sd.train(s, x1[train_idx], t[train_idx], epochs=args.epochs, batch=args.batch, lr=args.lr,
                          seed=args.seed, log=log)
        train_s = time.perf_counter() - t_start
        pred = s.forward(x1[hold])
        params = s.get_params()
    err = pred - t[hold]
    rmse = np.sqrt(np.mean(err ** 2, axis=0))
    hn = near[hold]

    def split(m):
        e = err[m]
        if not len(e):
            return None
        r = np.sqrt(np.mean(e ** 2, axis=0))
        return {"n": int(len(e)), "rmse_u_v_px": [float(r[0]), float(r[1])],
                "rmse_grad": [float(v / sd.GRAD_SCALE) for v in r[2:]],
                "p99_abs_u_px": float(np.percentile(np.abs(e[:, 0]), 99))}
    report = {"samples": int(len(x1)), "generated": int(generated), "held_out": int(n_hold),
              "epochs": args.epochs, "batch": args.batch, "lr": args.lr, "train_seconds": round(train_s, 2),
              "final_loss": losses[-1], "losses": [round(v, 6) for v in losses],
              "holdout_rmse_u_v_px": [float(rmse[0]), float(rmse[1])],
              "holdout_rmse_grad": [float(v / sd.GRAD_SCALE) for v in rmse[2:]],
              "holdout_p99_abs_u_px": float(np.percentile(np.abs(err[:, 0]), 99)),
              "holdout_near": split(hn), "holdout_far": split(~hn),
              "command": "python -m fasticgn_amd.seeder_train " + " ".join(sys.argv[1:])}
    log("[seeder] " + json.dumps(report))
    np.save(args.out, params)
    if args.report:
        Path(args.report).write_text(json.dumps(report, indent=1) + "\n")
    return 0


def _init(sd):
    """Fresh, seeded initialisation (not the shipped weights)."""
    rng = np.random.default_rng(20260)
    p = np.zeros(sd.param_count(), dtype=np.float32)
    o = 0
    for k, n in sd.LAYERS + [(128, sd.N_OUT)]:
        p[o:o + k * n] = rng.normal(0, np.sqrt(2.0 / k) * (1.0 if k * n != 128 * sd.N_OUT else 0.1), k * n)
        o += k * n + n
    return p


if __name__ == "__main__":
    sys.exit(main())
