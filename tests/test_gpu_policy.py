"""The device frame policy (through the C ABI) against the policy oracle.

1. The whole-frame fallback scenarios of smoke/run_frame_fallback_gate.py
   (tests/fallback_scenes.py) through the frame driver, eagerly and replayed
   as hipGraphs (the fallback is decided and run inside the graph): every
   frame is checked by one-step parity against the policy oracle seeded from
   the device's previous frame (same seeding, fallback and repair decisions,
   same counts, the parity gate), the gate's assertions 1-3, 5 and 6 hold on
   the device frames, and graph replay equals the eager frames byte for byte
   (the device's stand-in for the gate's thread-count independence).
2. Temporal extrapolation (--temporal-extrapolation) on a steadily moving
   sequence: one-step parity with the oracle's extrapolating seeds, the same
   number of extrapolated seeds per frame.
3. Output records packed on the device equal fi_format_raw of the downloaded
   POIs byte for byte, in template order (mixed radii included), also when
   every frame queues its records beside the next frame.
"""
import numpy as np
import pytest

import fasticgn_amd as fa
from fasticgn_amd import rawio, synth
from tests import fallback_scenes as fs
from tests.parity import compare
from tests.policy import PolicyOracle
from tests.test_policy import SEEDS

pytestmark = pytest.mark.gpu

KEYS = ("skip_fftcc_used", "fallback_used", "poi_fallback_count", "poi_fallback_recovered", "temporal_seed_count")


def say(*a):
    print(*a, flush=True)


def run_device(case, ref, tars, graph):
    out = []
    with fa.FrameDriver(case.radius) as d:
        d.set_graph(graph)
        d.set_reference(fa.Image2D(ref))
        d.set_pois(fs.template(case))
        for tar in tars:
            out.append(d.process(fa.Image2D(tar)))
    return out


def one_step_parity(oracle, case, ref, tars, dev, temporal=False, tag=""):
    """Frame k of the oracle seeded from the device's frame k-1; returns the
    oracle's stats (with the from-scratch pass of fallback frames)."""
    po = PolicyOracle(oracle, ref, fs.template(case), case.radius, temporal=temporal)
    ostats = []
    for k, (tar, (res, st)) in enumerate(zip(tars, dev)):
        o, ost = po.frame(tar, seed=dev[k - 1][0] if k else None)
        say(tag, "frame", k, {x: st[x] for x in KEYS}, {x: ost[x] for x in KEYS},
            "bad", st["pre_repair_bad_ratio"], ost["pre_repair_bad_ratio"])
        for key in KEYS:
            assert st[key] == ost[key], (k, key)
        cmp = compare(o, res, min_shared_valid=50)
        say(tag, "frame", k, {x: cmp.get(x) for x in ("status_mismatch", "p99_du", "p99_dv", "p99_dux", "max_du",
                                                      "iteration_mismatch")})
        assert cmp["pass"], (k, cmp)
        ostats.append(ost)
    return ostats


@pytest.mark.parametrize("name", ["cascade", "jump"])
def test_fallback_scenario_on_the_device(oracle, name):
    case = fs.CASES[name]
    ref, tars = fs.sequence(case, SEEDS[name])
    say("scenario", name, case.width, case.height, "r", case.radius, len(tars), "frames")
    eager = run_device(case, ref, tars, graph=False)
    graph = run_device(case, ref, tars, graph=True)
    for k, ((a, sa), (b, sb)) in enumerate(zip(eager, graph)):
        assert a.tobytes() == b.tobytes(), f"graph replay differs from the eager frame {k}"
        assert all(sa[x] == sb[x] for x in KEYS + ("bad_ratio", "valid_count"))
    ostats = one_step_parity(oracle, case, ref, tars, eager, tag=name)
    frames = [r for r, _ in eager]
    stats = [st for _, st in eager]
    errors, report = fs.check_case(case, frames, stats, [o.get("scratch") for o in ostats])
    for row in report:
        say(row)
    assert not errors, errors
    assert any(st["fallback_used"] for st in stats)


def test_temporal_extrapolation_matches_oracle(oracle):
    """A 512^2 scene moving 0.4 px per frame: from the third frame on the
    seeds are extrapolated, as many per frame as the oracle's."""
    steps = [(0.4 * k, -0.3 * k) for k in range(6)]
    base = synth.affine_scene("temporal", 512, 512, 6000, 20261205, t=(0.0, 0.0), grad=(0, 0, 0, 0),
                              radius=15, grid=12)
    tars = [synth.affine_scene("temporal", 512, 512, 6000, 20261205, t=t, grad=(2e-4 * k, 0, 0, 2e-4 * k),
                               radius=15, grid=12).tar for k, t in enumerate(steps)]
    ref = base.ref
    dev = []
    with fa.FrameDriver(15) as d:
        d.set_reference(fa.Image2D(ref))
        d.set_pois(base.pois)
        d.set_temporal_extrapolation(True)
        d.set_graph(True)
        for tar in tars:
            dev.append(d.process(fa.Image2D(tar)))
    po = PolicyOracle(oracle, ref, base.pois, 15, temporal=True)
    for k, (tar, (res, st)) in enumerate(zip(tars, dev)):
        o, ost = po.frame(tar, seed=dev[k - 1][0] if k else None)
        say("temporal frame", k, {x: st[x] for x in KEYS}, {x: ost[x] for x in KEYS})
        for key in KEYS:
            assert st[key] == ost[key], (k, key)
        cmp = compare(o, res)
        assert cmp["pass"], (k, cmp)
        if k >= 2:
            assert st["temporal_seed_count"] > 0.9 * len(base.pois)
            assert st["temporal_seed_used"] == 1
        else:
            assert st["temporal_seed_count"] == 0


def test_device_records_equal_formatted_pois(oracle):
    seq = synth.config2(size=256, frames=500)
    for mixed in (False, True):
        with fa.FrameDriver(15) as d:
            if mixed:
                d.set_self_adaptive(True)
            d.set_reference(fa.Image2D(seq.ref, depth12="right"))
            d.set_pois(seq.pois)
            for k in (0, 5):
                d.process(fa.Image2D(seq.frame(k), depth12="right"))
                res = d.download()
                rec = d.records()
                assert rawio.format_raw(res).tobytes() == rec.tobytes(), (mixed, k)
                want = np.zeros(len(res), dtype=rawio.RAW_DTYPE)
                want["x"], want["y"] = res["x"].astype(np.int32), res["y"].astype(np.int32)
                for f, j in (("u", 0), ("v", 6), ("ux", 1), ("uy", 2), ("vx", 7), ("vy", 8)):
                    want[f] = res["deformation"][:, j]
                want["zncc"] = res["result"][:, 2]
                assert rec.tobytes() == want.tobytes(), (mixed, k)
    # pipelined: every frame's records queued beside the next frame
    with fa.FrameDriver(15) as d:
        d.set_self_adaptive(True)
        d.set_graph(True)
        d.set_records(True)
        d.set_reference(fa.Image2D(seq.ref, depth12="right"))
        d.set_pois(seq.pois)
        d.upload_target(fa.Image2D(seq.frame(0), depth12="right"))
        last = None
        for k in range(1, 6):
            d.upload_target(fa.Image2D(seq.frame(k), depth12="right"))
            d.run()
            if last is not None:  # the frame before, its copy overlapped this frame
                assert d.records(1).tobytes() == last, k
            rec = d.records()
            res = d.download()
            last = rawio.format_raw(res).tobytes()
            assert rec.tobytes() == last, k
        with pytest.raises(fa.LogicError):
            d.reset_sequence()
            d.records()
