"""Multi-frame job from image files with the output-integrity ledger.

    python -m fasticgn_amd.run --ref REF --tar-list LIST --output-dir DIR
        [--radius 15] [--step 5] [--max-iteration 10] [--convergence 0.001]
        [--no-skip-fftcc] [--no-poi-fallback] [--temporal-extrapolation]
        [--learned-seed] [--input-depth compat8|12|16|auto]
        [--input-align auto|right|left] [--input-normalize none|to8|to16|unit]
        [--durability rename|fsync] [--repair] [--stop-after N]
    python -m fasticgn_amd.run --verify-only DIR

The multi-frame loop of the reference CLI (runMultiFrame,
opencorr_cli_cuda_mvp.cpp:3834-4509) reduced to the raw output format: one
reference, the targets of LIST (one image path per line; PGM or .npy), the
POI grid of buildPoiQueue (cli:2439-2472), every frame through the frame
driver, its 36-byte records packed on the device, published through the
ledger (fasticgn_amd/ledger.py), then the manifest and the completion
record.  Rerunning the same command resumes: a complete run is verified
and not recomputed; an interrupted one replays from the first frame and
keeps the payloads already published.  ``--verify-only`` audits a
directory without running anything.  One JSON status line goes to stdout;
the exit status is 0 on success.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
from pathlib import Path

import numpy as np

from . import ledger
from ._native import LIB_PATH
from .rawio import RAW_MANIFEST_NAME, RECORD_BYTES, RawWriter

OPTION_KEYS = ("radius", "step", "max_iteration", "convergence", "skip_fftcc", "poi_fallback",
               "temporal_extrapolation", "learned_seed", "format", "input_depth", "input_align", "input_normalize")


def _load(path: Path, depth: str, align: str, normalize: str):
    from .opencorr import Image2D
    if path.suffix == ".npy":
        px = np.load(path, allow_pickle=False)
        d = {"compat8": "native", "12": {"right": "right", "left": "left", "auto": "auto12"}[align], "16": "16",
             "auto": "auto" if px.dtype == np.uint16 else "native"}[depth]
        if depth == "compat8" and px.dtype == np.uint16:
            px = (px >> 8).astype(np.uint8)
        return Image2D(px, d, normalize)
    return Image2D.load(path, depth, align, normalize)


def _input(path: Path) -> dict:
    return {"path": str(path.resolve()), **ledger.fingerprint(path)}


def _targets(tar_list: Path) -> list[Path]:
    base = tar_list.parent
    out = []
    for line in tar_list.read_text(encoding="utf-8").splitlines():
        line = line.strip()
        if line and not line.startswith("#"):
            p = Path(line)
            out.append(p if p.is_absolute() else base / p)
    return out


def frame_file(k: int) -> str:
    return f"frame_{k:06d}.bin"


def layout(contract: dict, root: Path):
    """The files a contract promises: payloads, the manifest, the inputs."""
    root = Path(root)
    outs = [ledger.OutputSpec(key="frame/" + t["file"], path=root / t["file"], record_bytes=RECORD_BYTES)
            for t in contract["targets"]]
    inputs = [(contract["reference"]["path"], {k: contract["reference"][k] for k in ("bytes", "digest")})]
    inputs += [(t["input"]["path"], {k: t["input"][k] for k in ("bytes", "digest")}) for t in contract["targets"]]
    return outs, [root / RAW_MANIFEST_NAME], inputs


def run(args) -> dict:
    from .driver import DEFAULT_POLICY, LEARNED_SEED, POI_FALLBACK, SKIP_FFTCC, FrameDriver
    from .opencorr import grid_queue
    out_dir = Path(args.output_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    ref_path, tar_list = Path(args.ref), Path(args.tar_list)
    targets = _targets(tar_list)
    if not targets:
        raise ledger.IntegrityError("the target list is empty")
    options = {"radius": args.radius, "step": args.step, "max_iteration": args.max_iteration,
               "convergence": args.convergence, "skip_fftcc": not args.no_skip_fftcc,
               "poi_fallback": not args.no_poi_fallback, "temporal_extrapolation": args.temporal_extrapolation,
               "learned_seed": args.learned_seed, "format": "raw", "input_depth": args.input_depth,
               "input_align": args.input_align, "input_normalize": args.input_normalize}
    tlist = [{"input": _input(p), "output_name": p.name, "file": frame_file(k)} for k, p in enumerate(targets)]
    contract = ledger.make_contract({"library": LIB_PATH.name, **ledger.fingerprint(LIB_PATH)}, _input(ref_path),
                                    options, tlist)
    outputs, metadata, _ = layout(contract, out_dir)
    allowed = (tar_list.name,) if tar_list.resolve().parent == out_dir.resolve() else ()
    status = {"status": "ok", "output_dir": str(out_dir), "frames": len(targets), "published": 0, "kept": 0,
              "durability": args.durability}
    with ledger.OutputSession(out_dir, contract, outputs, metadata, args.durability, args.repair, allowed) as s:
        s.begin()
        status.update(verified_files=s.verified_files, verified_bytes=s.verified_bytes,
                      repaired_files=s.repaired_files, archived_residue=s.archived_residue,
                      quarantine=str(s.quarantine_directory) if s.quarantine_directory else None)
        if s.complete:
            status["status"] = "complete"
            return status
        policy = DEFAULT_POLICY
        if args.no_skip_fftcc:
            policy &= ~SKIP_FFTCC
        if args.no_poi_fallback:
            policy &= ~POI_FALLBACK
        seeder = None
        if args.learned_seed:
            from .seeder import LearnedSeeder
            seeder = LearnedSeeder()
            policy = LEARNED_SEED
        ref = _load(ref_path, args.input_depth, args.input_align, args.input_normalize)
        h, w = ref.data.shape
        writer = RawWriter(out_dir, ref_path=str(ref_path), tar_list_path=str(tar_list), subset_radius=args.radius,
                           grid_space=args.step, conv_threshold=args.convergence,
                           max_iteration=args.max_iteration, policy=policy, learned_seed=args.learned_seed,
                           temporal_extrapolation=args.temporal_extrapolation, session=s)
        writer.head["durability"] = args.durability
        with FrameDriver(args.radius, args.convergence, args.max_iteration) as d:
            d.set_records(False)
            if seeder is not None:
                d.set_seeder(seeder)
            if args.temporal_extrapolation:
                d.set_temporal_extrapolation(True)
            d.set_reference(ref)
            d.set_pois(grid_queue(w, h, args.radius, args.step))
            for k, (tp, spec) in enumerate(zip(targets, outputs)):
                if args.stop_after is not None and k >= args.stop_after:
                    status["status"] = "stopped"
                    return status
                _, qa = d.process(_load(tp, args.input_depth, args.input_align, args.input_normalize), policy)
                what = writer.write_frame(d.records(), qa, tar_path=str(tp), output_name=spec.path.name)["ledger"]
                status[what] += 1
        writer.close()
        s.finish()
        if seeder is not None:
            seeder.close()
    return status


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m fasticgn_amd.run", description=__doc__.split("\n\n")[0])
    ap.add_argument("--verify-only", metavar="DIR")
    ap.add_argument("--ref")
    ap.add_argument("--tar-list")
    ap.add_argument("--output-dir")
    ap.add_argument("--radius", type=int, default=15)
    ap.add_argument("--step", type=int, default=5)
    ap.add_argument("--max-iteration", type=int, default=10)
    ap.add_argument("--convergence", type=float, default=0.001)
    ap.add_argument("--no-skip-fftcc", action="store_true")
    ap.add_argument("--no-poi-fallback", action="store_true")
    ap.add_argument("--temporal-extrapolation", action="store_true")
    ap.add_argument("--learned-seed", action="store_true")
    ap.add_argument("--input-depth", default="compat8", choices=["compat8", "12", "16", "auto"])
    ap.add_argument("--input-align", default="auto", choices=["auto", "right", "left"])
    ap.add_argument("--input-normalize", default="none", choices=["none", "to8", "to16", "unit"])
    ap.add_argument("--durability", default="rename", choices=["rename", "fsync"])
    ap.add_argument("--repair", action="store_true")
    ap.add_argument("--stop-after", type=int, default=None,
                    help="publish at most N frames and stop (the next run resumes)")
    args = ap.parse_args(argv)
    if args.verify_only is not None:
        rep = ledger.audit_run_directory(args.verify_only, layout)
        print(json.dumps(rep), flush=True)
        if rep["status"] != "complete":
            print(f"verify-only: {rep['status']} ({args.verify_only})", file=sys.stderr)
        return rep["exit_code"]
    if not (args.ref and args.tar_list and args.output_dir):
        ap.error("--ref, --tar-list and --output-dir are required (or --verify-only DIR)")
    try:
        status = run(args)
    except ledger.IntegrityError as e:
        print(json.dumps({"status": "error", "error": str(e)}), flush=True)
        print(f"Error: {e}", file=sys.stderr)
        return 1
    print(json.dumps(status), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
