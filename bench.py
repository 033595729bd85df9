"""Throughput of the 2D-DIC hot path on one MI355X.

Default workload (one step = one frame of the hot path against a resident
reference): 2048x2048 u8 speckle, subset radius 15 (31x31), POIs on a 5 px grid
with margin 17 (162,409 POIs) starting from the global integer guess, ICGN2D1
(conv 1e-3, at most 10 iterations).  The reference image, its gradients, the
FFTCC reference spectra and the per-POI Hessians are prepared once before
timing; the raw target is resident in HBM.  A step runs the target B-spline
table, the seeding (FFTCC, or the learned seeder with its FFTCC fallback), IC-GN
and the frame QA, and leaves the POIs in HBM.  Both seeding modes are timed.

Prints one JSON line.  value = converged POIs per second (zncc finite >= 0) of
the learned-seed path when it meets the p99 <= 0.01 px gate against the CPU
oracle (else of the FFTCC path), frames replayed as captured hipGraphs (each
mode is also timed eagerly, which gives the per-stage times, and the graph
results are checked byte for byte against the eager ones); detail holds both
modes, their ratio, the error against the known field and the parity against
the FP32 and FP64 oracles.
roofline: the IC-GN kernel, algorithmic bytes (2r+1)^2 * 76 B per executed
POI-iteration over its HIP-event duration on the driver's stream.
cpu_baseline: the CPU oracle (FP32 restatement of OpenCorr FFTCC2D + ICGN2D1,
OpenMP) on a bounded sample of the same frame.
--workload config2 / config4: the multi-frame sequence and the crack field.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "converged POIs/sec (31x31 subset), learned vs FFTCC seed at p99 <=0.01 px"
HBM_PEAK_GBS = 8000.0
BYTES_PER_SAMPLE = 76  # 64 B B-spline texel + 4 B reference + 8 B gradients


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profile(name: str):
    """HBM bytes per IC-GN launch from a committed rocprofv3 --pmc summary
    (profiles/<name>.json), or None."""
    p = ROOT / "profiles" / f"{name}.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("hbm_bytes_per_dispatch")
    except Exception:
        return None


def cpu_baseline(sc, sample_every: int):
    """Time the CPU oracle on every sample_every-th POI: FFTCC (uncached, as
    opencorr_cli.cpp times FFTCC2D::compute) + ICGN2D1 prepare (gradients of the
    reference, B-spline table of the target: whole images) + compute.  Scale the
    per-POI part to the full frame."""
    from oracle import pyoracle

    threads = min(16, os.cpu_count() or 1)
    os.environ["ORACLE_THREADS"] = str(threads)
    ref32, tar32 = sc.ref.astype(np.float32), sc.tar.astype(np.float32)
    pick = np.arange(0, len(sc.pois), sample_every)
    q = sc.pois[pick].copy()
    t0 = time.perf_counter()
    pyoracle.fftcc(ref32, tar32, sc.radius, q)
    t1 = time.perf_counter()
    gx, gy = pyoracle.gradient(ref32)
    coef = pyoracle.bspline(tar32)
    t2 = time.perf_counter()
    pyoracle.icgn(ref32, tar32, sc.radius, q, prepared=(gx, gy, coef))
    t3 = time.perf_counter()
    per_poi = ((t1 - t0) + (t3 - t2)) / len(pick)
    t_full = (t2 - t1) + per_poi * len(sc.pois)
    valid_frac = float(np.mean(np.isfinite(q["result"][:, 2]) & (q["result"][:, 2] >= 0)))
    value = valid_frac * len(sc.pois) / t_full
    return {
        "value": value, "unit": "POIs/s", "cores": threads, "kind": "port",
        "sample": (f"{len(pick)} of {len(sc.pois)} POIs (every {sample_every}th): FFTCC2D uncached "
                   f"{t1 - t0:.2f} s + ICGN2D1 prepare {t2 - t1:.2f} s (whole images) + compute "
                   f"{t3 - t2:.2f} s; per-POI part scaled to the frame"),
    }, q, pick


def copy_calibration(mib: int, repeats: int = 3) -> int:
    """Launch the HIP runtime's device-to-device copy kernel on a buffer of
    known size (reads mib MiB, writes mib MiB): the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run."""
    import ctypes
    import fasticgn_amd  # noqa: F401  (same HIP runtime as the library)
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    for _ in range(repeats):
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": n, "repeats": repeats}), flush=True)
    return 0


def tune_summary(rows: list) -> dict:
    """Best autotuner time per (radius class, launch kind); radius 0 = one
    launch over every class."""
    out = {}
    for e in rows:
        key = f"r{e['radius'] or 'all'}_{'queue' if e['persistent'] else 'blocks'}"
        if key not in out or e["ms"] < out[key]:
            out[key] = round(e["ms"], 4)
    return out


_C2 = None


def _c2_frame(job):
    """Render frame k of the config-2 sequence (process-pool worker)."""
    global _C2
    size, frames, k = job
    if _C2 is None:
        from fasticgn_amd import synth
        _C2 = synth.config2(size=size, frames=frames)
    return _C2.frame(k)


def bench_config2(args) -> int:
    """500 frames of the mixed-radius u16 sequence through the frame driver:
    hipGraph per frame, autotuned launches, default policy (seed from the
    previous frame, bad-ratio guard, POI-level repair).  value = converged
    POIs per second over the timed frames, each frame's target uploaded from
    host memory inside the timed region (the sequence streams in; target k+1
    is staged and copied while frame k runs)."""
    from concurrent.futures import ProcessPoolExecutor
    import fasticgn_amd as fa
    from fasticgn_amd import synth
    from fasticgn_amd._native import synchronize as sync

    size, F = 1024, args.frames
    sq = synth.config2(size=size, frames=F)
    t0 = time.perf_counter()
    with ProcessPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        frames = list(ex.map(_c2_frame, [(size, F, k) for k in range(F)], chunksize=8))
    log(f"[bench] config2: {F} frames of {size}^2 u16 rendered in {time.perf_counter() - t0:.1f} s, "
        f"{len(sq.pois)} POIs, radii {sq.meta['radii']}")
    img = lambda a: fa.Image2D(a, depth12="right")  # noqa: E731
    drv = fa.FrameDriver(15, 0.001, 10)
    drv.set_self_adaptive(True)
    drv.set_reference(img(sq.ref))
    drv.set_pois(sq.pois)
    drv.set_graph(True)
    drv.upload_target(img(frames[F // 2]))
    tune = drv.autotune()
    chosen = [e for e in tune if e["chosen"]]
    log(f"[bench] autotune: {len(tune)} configurations, chosen {chosen}")
    drv.reset_sequence()
    W = max(1, min(args.warmup, F - 1))
    for k in range(W):
        drv.process(img(frames[k]))
    sync()
    stats = []
    t0 = time.perf_counter()
    drv.upload_target(img(frames[W]))
    for k in range(W, F):
        if k + 1 < F:
            drv.upload_target(img(frames[k + 1]))  # staged while frame k runs
        stats.append(drv.run())
    sync()
    wall = time.perf_counter() - t0
    res = drv.download()
    K = F - W
    valid = float(sum(s["valid_count"] for s in stats))
    dev_ms = [s["device_ms"] for s in stats]
    u, v, *_ = sq.truth(F - 1, res["x"], res["y"])
    ok = (res["result"][:, 2] >= 0) & np.isfinite(res["result"][:, 2])
    r = sq.pois["subset_radius"][:, 0].astype(int)
    acc = {}
    for rad in sq.meta["radii"]:
        m = ok & (r == rad)
        e = np.maximum(np.abs(res["deformation"][m, 0] - u[m]), np.abs(res["deformation"][m, 6] - v[m]))
        acc[str(rad)] = {"valid": int(m.sum()), "of": int((r == rad).sum()),
                         "p99_err_px": round(float(np.percentile(e, 99)), 5), "max_err_px": round(float(e.max()), 5)}
    out = {
        "metric": "converged POIs/sec, mixed radii 10/15/20, 500-frame streaming sequence",
        "value": round(valid / wall, 1), "unit": "POIs/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(wall * 1e3 / K, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{size}x{size} u16 12-bit speckle, {F} frames (t = 15 k/499 px at 30 deg, "
                               f"strain 5e-3 k/499), {len(sq.pois)} POIs on a 5 px grid, radii 10/15/20 "
                               "(POI index mod 3), seeded from the previous frame, hipGraph per frame, autotuned",
                   "pois": len(sq.pois), "parallelism": "single-gpu"},
        "detail": {"device_ms_median": round(float(np.median(dev_ms)), 4),
                   "device_only_pois_per_s": round(valid / (sum(dev_ms) / 1e3), 1),
                   "upload_inclusive": True,
                   "iterations_per_poi": round(float(np.mean([s["iteration_mean"] for s in stats])), 3),
                   "poi_fallback_total": int(sum(s["poi_fallback_count"] for s in stats)),
                   "frame_fallbacks": int(sum(s["fallback_used"] for s in stats)),
                   "autotune_chosen": chosen, "autotune_best_ms": tune_summary(tune),
                   "last_frame_accuracy": acc},
    }
    print(json.dumps(out), flush=True)
    drv.close()
    return 0


def bench_config4(args) -> int:
    """The 4096^2 crack field, 1.03 M POIs at r = 20, one frame: learned seed
    with FFTCC fallback (and the FFTCC-seeded path beside it), autotuned
    IC-GN launch.  Reports accuracy against the truth away from the crack,
    parity against the FP64 oracle on a POI sample, and the fraction of POIs
    within 2r of the crack sent to FFTCC."""
    import fasticgn_amd as fa
    from fasticgn_amd import seeder as sd
    from fasticgn_amd import synth
    from fasticgn_amd._native import synchronize as sync

    t0 = time.perf_counter()
    sc = synth.config4()
    log(f"[bench] config4: {sc.ref.shape} u16, {len(sc.pois)} POIs ({time.perf_counter() - t0:.1f} s)")
    drv = fa.FrameDriver(sc.radius, 0.001, 10)
    drv.set_reference(fa.Image2D(sc.ref))
    drv.set_pois(sc.pois)
    drv.upload_target(fa.Image2D(sc.tar))
    drv.run(policy=0)
    tune = drv.autotune()
    chosen = [e for e in tune if e["chosen"]]
    log(f"[bench] autotune chosen {chosen}")
    seeder = sd.LearnedSeeder()
    drv.set_seeder(seeder, args.seed_zncc)
    # POIs start at guess 0 with u up to ~6 px: a second seeder pass crops the
    # target at the first pass's rounded u, v (the seeder's near regime)
    drv.set_seed_passes(2)
    runs = {}
    steps = max(1, min(args.steps, 10))
    for name, pol in (("fftcc", 0), ("learned", fa.driver.LEARNED_SEED)):
        drv.run(policy=pol)
        sync()
        st = []
        t0 = time.perf_counter()
        for _ in range(steps):
            st.append(drv.run(policy=pol))
        sync()
        wall = time.perf_counter() - t0
        runs[name] = {"ms": wall * 1e3 / steps, "stats": st[-1], "res": drv.download(), "redo": drv.download_redo()}
        log(f"[bench] {name}: {runs[name]['ms']:.3f} ms/frame, valid {st[-1]['valid_count']}")
    r = runs["learned"]
    res = r["res"]
    dist = sc.crack_distance(res["x"], res["y"])
    near = dist <= 2 * sc.radius
    ok = (res["result"][:, 2] >= 0) & np.isfinite(res["result"][:, 2])
    u, *_ = sc.truth(res["x"], res["y"])
    far = ok & ~near
    err = np.abs(res["deformation"][far, 0] - u[far])
    errv = np.abs(res["deformation"][far, 6])
    # FP64 oracle on a sample away from the crack, FFTCC-seeded
    from oracle import pyoracle
    from tests.parity import compare
    rng = np.random.default_rng(4)
    pick = np.sort(rng.choice(np.nonzero(~near)[0], 2000, replace=False))
    ref32 = pyoracle.as_f32(sc.ref)
    tar32 = pyoracle.as_f32(sc.tar)
    o = sc.pois[pick].copy()
    pyoracle.fftcc(ref32, tar32, sc.radius, o)
    pyoracle.icgn(ref32, tar32, sc.radius, o, f64=True)
    par = compare(o, res[pick])
    keys = ("n", "status_mismatch", "median_du", "p99_du", "max_du", "median_dv", "p99_dv", "max_dv", "pass")
    st = r["stats"]
    value = st["valid_count"] / (r["ms"] / 1e3)
    out = {
        "metric": "converged POIs/sec (41x41 subset), learned seed with FFTCC fallback",
        "value": round(value, 1), "unit": "POIs/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": round(r["ms"], 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"4096x4096 u16 speckle (noise 2), crack at 30 deg through the centre with a 1.5 px "
                               f"jump of u, u = 1e-3 x; grid 4, r = 20, {len(sc.pois)} POIs; learned seed (2 passes) "
                               "+ FFTCC fallback + ICGN2D1 conv 1e-3 max 10 it, autotuned",
                   "pois": len(sc.pois), "parallelism": "single-gpu"},
        "detail": {"fftcc_path_pois_per_s": round(runs["fftcc"]["stats"]["valid_count"] / (runs["fftcc"]["ms"] / 1e3), 1),
                   "learned_vs_fftcc": round(value / (runs["fftcc"]["stats"]["valid_count"] / (runs["fftcc"]["ms"] / 1e3)), 4),
                   "seed_fallback_fraction": round(st["seed_fallback_count"] / len(sc.pois), 5),
                   "near_crack_pois": int(near.sum()),
                   "near_crack_fallback_fraction": round(float(r["redo"][near].mean()), 5),
                   "far_fallback_fraction": round(float(r["redo"][~near].mean()), 5),
                   "valid_fraction_near_crack": round(float(ok[near].mean()), 5),
                   "valid_fraction_far": round(float(ok[~near].mean()), 5),
                   "truth_far_p99_du_px": round(float(np.percentile(err, 99)), 5),
                   "truth_far_p99_dv_px": round(float(np.percentile(errv, 99)), 5),
                   "iterations_per_poi": round(st["iteration_sum"] / len(sc.pois), 3),
                   "seed_ms": round(st["seed_ms"], 4), "icgn_ms": round(st["icgn_ms"], 4),
                   "device_ms": round(st["device_ms"], 4),
                   "parity_vs_fp64_oracle_far_sample": {k: par.get(k) for k in keys},
                   "autotune_chosen": chosen, "autotune_best_ms": tune_summary(tune)},
    }
    print(json.dumps(out), flush=True)
    drv.close()
    seeder.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--cpu-sample-every", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--shape", default="", help="IC-GN lane-to-POI shape WAVES,ROWS")
    ap.add_argument("--sampling", default="tune", choices=["tune", "auto", "table", "group"],
                    help="IC-GN launch: 'tune' = the autotuner's choice on this frame (lane shape, sampling, "
                         "launch mode), or a fixed sampling mode (per-POI LDS tile, table only, or tiles "
                         "shared by four neighbouring POIs)")
    ap.add_argument("--seeder", default="auto",
                    help="learned seeder weights (.npy), 'auto' = the shipped weights, 'off' = FFTCC path only")
    ap.add_argument("--seed-zncc", type=float, default=0.8,
                    help="ZNCC below which a learned-seed POI is redone from FFTCC")
    ap.add_argument("--workload", default="config1", choices=["config1", "config2", "config4"],
                    help="config1: the 2048^2 single pair (default); config2: the 500-frame mixed-radius "
                         "u16 sequence; config4: the 4096^2 crack field (1.03 M POIs, r = 20)")
    ap.add_argument("--frames", type=int, default=500, help="config2: frames of the sequence")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported by this build "
              "(the path would shard over POIs: every GPU holding its shard's reference-side state and "
              "receiving each frame's target)", file=sys.stderr)
        return 2
    if args.workload == "config2":
        return bench_config2(args)
    if args.workload == "config4":
        return bench_config4(args)

    import fasticgn_amd as fa
    from fasticgn_amd import synth
    from fasticgn_amd._native import synchronize as sync

    t0 = time.perf_counter()
    sc = synth.config1(size=args.size)
    log(f"[bench] scene {sc.name}: {sc.ref.shape} {len(sc.pois)} POIs, guess {sc.meta['initial_guess']}"
        f" ({time.perf_counter() - t0:.1f} s)")

    drv = fa.FrameDriver(sc.radius, 0.001, 10)
    drv.set_reference(fa.Image2D(sc.ref))
    drv.set_pois(sc.pois)
    if args.shape:
        drv.set_shape(*[int(v) for v in args.shape.split(",")])
    if args.sampling != "tune":
        drv.set_sampling(args.sampling)
    t0 = time.perf_counter()
    drv.upload_target(fa.Image2D(sc.tar))
    upload_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    drv.run(policy=0)  # first run also prepares the reference side
    cold_s = time.perf_counter() - t0
    log(f"[bench] cold frame (target upload + reference prep + frame) {(upload_s + cold_s) * 1e3:.1f} ms")
    tuned = None
    if args.sampling == "tune":
        # untimed: times every launch configuration on this frame's FFTCC
        # seeds and keeps the fastest for the frames below
        drv.upload_target(fa.Image2D(sc.tar))
        rows = drv.autotune()
        tuned = [e for e in rows if e["chosen"]][0]
        log(f"[bench] autotune: {len(rows)} configurations, chosen {tuned}")

    def timed_once(policy: int, label: str, graph: bool) -> dict:
        drv.set_graph(graph)
        for _ in range(args.warmup):
            drv.run(policy=policy)
        sync()
        stats = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            stats.append(drv.run(policy=policy))
        sync()
        wall = time.perf_counter() - t0
        ms = wall * 1e3 / args.steps
        mean = {k: float(np.mean([s[k] for s in stats])) for k in stats[0]}
        conv = mean["valid_count"]
        log(f"[bench] {label}{' (graph)' if graph else ''}: {args.steps} steps in {wall:.3f} s, "
            f"{conv / (ms / 1e3):.4g} POIs/s, {mean['iteration_sum'] / len(sc.pois):.3f} it/POI")
        return {"ms": ms, "value": conv / (ms / 1e3), "mean": mean, "res": drv.download()}

    def timed(policy: int, label: str) -> dict:
        """Eager frames (per-stage HIP-event times: the roofline's kernel
        time), then the same frames replayed as captured graphs (throughput)."""
        eager = timed_once(policy, label, False)
        graph = timed_once(policy, label, True)
        if graph["res"].tobytes() != eager["res"].tobytes():
            raise RuntimeError(f"{label}: graph replay differs from the eager frame")
        out = dict(graph)
        out["eager"] = eager
        # stage times (NaN in graph replay) from the eager frames
        out["mean"] = dict(graph["mean"])
        for k in ("bspline_ms", "fftcc_ms", "icgn_ms", "seed_ms"):
            out["mean"][k] = eager["mean"][k]
        return out

    runs = {"fftcc": timed(0, "FFTCC seed")}
    seeder = None
    if args.seeder != "off":
        from fasticgn_amd import seeder as sd
        wpath = Path(args.seeder) if args.seeder != "auto" else sd.WEIGHTS_PATH
        if wpath.exists():
            seeder = sd.LearnedSeeder(np.load(wpath))
            drv.set_seeder(seeder, args.seed_zncc)
            runs["learned"] = timed(fa.driver.LEARNED_SEED, "learned seed")
        else:
            log(f"[bench] no seeder weights at {wpath}: FFTCC path only")

    cpu = None
    parity, parity64 = {}, {}
    # error against the known displacement field, every POI
    truth = {}
    for name, r in runs.items():
        res = r["res"]
        ok = np.isfinite(res["result"][:, 2]) & (res["result"][:, 2] >= 0)
        u, v, *_ = sc.truth(res["x"], res["y"])
        e = np.maximum(np.abs(res["deformation"][ok, 0] - u[ok]), np.abs(res["deformation"][ok, 6] - v[ok]))
        truth[name] = {"median_px": round(float(np.median(e)), 6), "p99_px": round(float(np.percentile(e, 99)), 6),
                       "max_px": round(float(e.max()), 6)}
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle) ...")
        cpu, oracle_q, pick = cpu_baseline(sc, args.cpu_sample_every)
        from tests.parity import compare
        keys = ("n", "status_mismatch", "median_du", "p99_du", "max_du", "median_dv", "p99_dv", "max_dv", "pass")
        for name, r in runs.items():
            par = compare(oracle_q, r["res"][pick])
            parity[name] = {k: par.get(k) for k in keys}
            log(f"[bench] parity {name}: {parity[name]}")
        log(f"[bench] CPU baseline {cpu['value']:.0f} POIs/s")
        # FP64 twin of the oracle on every 16th POI (FFTCC-seeded)
        from oracle import pyoracle
        pick64 = np.arange(0, len(sc.pois), 16)
        o64 = sc.pois[pick64].copy()
        ref32, tar32 = sc.ref.astype(np.float32), sc.tar.astype(np.float32)
        pyoracle.fftcc(ref32, tar32, sc.radius, o64)
        pyoracle.icgn(ref32, tar32, sc.radius, o64, f64=True)
        for name, r in runs.items():
            par = compare(o64, r["res"][pick64])
            parity64[name] = {k: par.get(k) for k in keys}
        log(f"[bench] FP64-oracle parity: {parity64}")

    def gate(name: str) -> bool:
        """p99 <= 0.01 px (and median, max) against the CPU oracle."""
        par = parity.get(name)
        if par is None:
            return True  # not checked in this run
        return (max(par["p99_du"], par["p99_dv"]) <= 1e-2 and max(par["median_du"], par["median_dv"]) <= 1e-3
                and max(par["max_du"], par["max_dv"]) <= 0.5)

    chosen = "learned" if "learned" in runs and gate("learned") else "fftcc"

    # the same frames with the target handed over in host memory every frame:
    # target k+1 is queued (pinned staging + DMA on the copy stream) while
    # frame k runs
    pol = fa.driver.LEARNED_SEED if chosen == "learned" else 0
    tar_img = fa.Image2D(sc.tar)
    drv.upload_target(tar_img)
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        drv.upload_target(tar_img)
        drv.run(policy=pol)
    sync()
    pcie_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    drv.run(policy=pol)  # drain the queued target
    r = runs[chosen]
    m = r["mean"]
    ns = (2 * sc.radius + 1) ** 2
    alg_bytes = m["iteration_sum"] * ns * BYTES_PER_SAMPLE
    achieved = alg_bytes / (m["icgn_ms"] / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile("icgn_config1"),
                "kernel": (("k_icgn_group" if tuned["sampling"] == 2 else "k_icgn")
                           + f"<{tuned['waves']},{tuned['rows_per_band']}>" if tuned
                           else f"k_icgn<{args.shape or 'default'}, {args.sampling}>"),
                "kernel_ms": round(m["icgn_ms"], 4),
                "algorithmic_bytes": int(alg_bytes)}

    def detail(name: str) -> dict:
        d = runs[name]
        mm = d["mean"]
        out = {"pois_per_s": round(d["value"], 1), "ms_per_step": round(d["ms"], 4),
               "eager_pois_per_s": round(d["eager"]["value"], 1), "eager_ms_per_step": round(d["eager"]["ms"], 4),
               "device_ms": round(mm["device_ms"], 4), "bspline_ms": round(mm["bspline_ms"], 4),
               "fftcc_ms": round(mm["fftcc_ms"], 4), "icgn_ms": round(mm["icgn_ms"], 4),
               "iterations_per_poi": round(mm["iteration_sum"] / len(sc.pois), 3),
               "converged_per_step": mm["valid_count"], "parity_vs_cpu_oracle": parity.get(name),
               "parity_vs_fp64_oracle": parity64.get(name), "error_vs_truth": truth[name]}
        if name == "learned":
            out.update({"seed_ms": round(mm["seed_ms"], 4), "fftcc_fallback_pois": mm["seed_fallback_count"],
                        "fftcc_fallback_ms": round(mm["fftcc_ms"], 4), "zncc_threshold": args.seed_zncc})
        return out

    seeding = {"fftcc": "FFTCC seed from the global integer guess",
               "learned": "learned seed (BF16-MFMA network, FFTCC fallback below the ZNCC threshold)"}[chosen]
    out = {
        "metric": METRIC, "value": round(r["value"], 1), "unit": "POIs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r["ms"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"2048x2048 u8 speckle (seeded), r=15 (31x31), grid 5, {len(sc.pois)} POIs, "
                               f"{seeding} + ICGN2D1 conv 1e-3 max 10 it, reference resident, "
                               + ("autotuned IC-GN launch" if tuned else f"IC-GN sampling {args.sampling}"),
                   "seeder": chosen, "pois": len(sc.pois), "parallelism": "single-gpu"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": {name: detail(name) for name in runs},
    }
    out["detail"].update({"cold_frame_ms": round((upload_s + cold_s) * 1e3, 2),
                          "pcie_inclusive_ms_per_step": round(pcie_ms, 4),
                          "pcie_inclusive_pois_per_s": round(m["valid_count"] / (pcie_ms / 1e3), 1)})
    if "learned" in runs:
        out["detail"]["learned_vs_fftcc"] = round(runs["learned"]["value"] / runs["fftcc"]["value"], 4)
    if tuned:
        out["detail"]["autotune_chosen"] = tuned
    print(json.dumps(out), flush=True)
    drv.close()
    if seeder is not None:
        seeder.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
