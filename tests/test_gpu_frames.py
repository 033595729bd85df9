"""Batched multi-frame path on the device (through the C ABI): mixed subset
radii (one radius class per radius), frames replayed as captured hipGraphs,
the persistent work-queue IC-GN launch and the autotuner.

Bars: the first frame of every radius class meets the parity gate against the
CPU oracle run with that radius (validity flags identical; u, v median
<= 1e-3, p99 <= 1e-2, max <= 0.5 px); a mixed-radius sequence is
byte-identical to one single-radius driver per class; graph replay and the
persistent launch are byte-identical to the eager, one-block-per-POI frames;
the tuned configuration meets the parity gate against the default one.
"""
import numpy as np
import pytest

import fasticgn_amd as fa
from fasticgn_amd import synth
from tests.parity import compare

pytestmark = pytest.mark.gpu

STEPS = [0, 8, 16, 24, 32, 40]  # frames of the 500-frame motion (0.24 px apart)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def seq():
    return synth.config2(size=512, frames=500)


@pytest.fixture(scope="module")
def targets(seq):
    return {k: seq.frame(k) for k in STEPS}


def img(a):
    return fa.Image2D(a, depth12="right")


def mixed_driver(seq, **kw):
    d = fa.FrameDriver(15)
    d.set_self_adaptive(True)
    d.set_reference(img(seq.ref))
    d.set_pois(seq.pois)
    if kw.get("graph"):
        d.set_graph(True)
    if kw.get("persistent"):
        d.set_persistent(kw["persistent"])
    return d


def run_sequence(d, targets, policy=fa.driver.DEFAULT_POLICY):
    out = []
    for k in STEPS:
        res, st = d.process(img(targets[k]), policy=policy)
        out.append((res, st))
    return out


def same_bytes(a, b):
    return a.tobytes() == b.tobytes()


def test_mixed_radii_first_frame_matches_oracle(oracle, seq, targets):
    k = STEPS[-1]
    with mixed_driver(seq) as d:
        res, st = d.process(img(targets[k]), policy=0)
    say("mixed first frame", {x: st[x] for x in ("valid_count", "iteration_mean", "fftcc_pois")})
    assert st["fftcc_pois"] == len(seq.pois)
    ref32 = oracle.as_f32(seq.ref, "right")
    tar32 = oracle.as_f32(targets[k], "right")
    r = seq.pois["subset_radius"][:, 0].astype(int)
    assert np.array_equal(res["subset_radius"][:, 0].astype(int), r)
    for rad in synth.RADII_MIXED:
        idx = np.nonzero(r == rad)[0][::3]  # a third of each class
        o = seq.pois[idx].copy()
        oracle.fftcc(ref32, tar32, rad, o)
        g = res[idx]
        # FFTCC integer seeds are bit-exact: IC-GN starts from them (result[0:2])
        assert np.array_equal(o["deformation"][:, 0], g["result"][:, 0])
        assert np.array_equal(o["deformation"][:, 6], g["result"][:, 1])
        oracle.icgn(ref32, tar32, rad, o)
        cmp = compare(o, g)
        say("radius", rad, {x: cmp.get(x) for x in ("n", "status_mismatch", "p99_du", "p99_dv", "max_du")})
        assert cmp["pass"], cmp
    u, v, *_ = seq.truth(k, res["x"], res["y"])
    ok = res["result"][:, 2] >= 0
    err = np.maximum(np.abs(res["deformation"][ok, 0] - u[ok]), np.abs(res["deformation"][ok, 6] - v[ok]))
    say("truth p99", float(np.percentile(err, 99)))
    assert ok.mean() > 0.99 and np.percentile(err, 99) < 0.05


def test_mixed_radii_sequence_matches_per_class_drivers(seq, targets):
    with mixed_driver(seq) as d:
        mixed = run_sequence(d, targets)
    r = seq.pois["subset_radius"][:, 0].astype(int)
    for rad in synth.RADII_MIXED:
        idx = np.nonzero(r == rad)[0]
        with fa.FrameDriver(rad) as d:
            d.set_reference(img(seq.ref))
            d.set_pois(seq.pois[idx].copy())
            single = run_sequence(d, targets)
        for f, ((mres, mst), (sres, sst)) in enumerate(zip(mixed, single)):
            assert same_bytes(mres[idx], sres), (rad, f)
            assert mst["skip_fftcc_used"] == sst["skip_fftcc_used"]
            assert not mst["fallback_used"] and not sst["fallback_used"]
    say("per-frame", [{x: st[x] for x in ("skip_fftcc_used", "poi_fallback_count", "iteration_mean",
                                           "valid_ratio")} for _, st in mixed])


def test_graph_replay_matches_eager(seq, targets):
    with mixed_driver(seq) as d:
        eager = run_sequence(d, targets)
    with mixed_driver(seq, graph=True) as d:
        graph = run_sequence(d, targets)
        # the graphs replay again after a reset, with eager frames between
        d.reset_sequence()
        again = run_sequence(d, targets)
    keys = ("skip_fftcc_used", "fallback_used", "poi_fallback_count", "poi_fallback_recovered", "valid_count",
            "iteration_sum", "iteration_max", "fftcc_pois", "bad_ratio")
    for f, ((er, es), (gr, gs), (ar, as_)) in enumerate(zip(eager, graph, again)):
        assert same_bytes(er, gr) and same_bytes(er, ar), f
        for key in keys:
            assert es[key] == gs[key] == as_[key], (f, key)
        assert np.isnan(gs["icgn_ms"]) and gs["device_ms"] > 0
    say("graph frames", [round(st["device_ms"], 3) for _, st in graph], "eager",
        [round(st["device_ms"], 3) for _, st in eager])


def test_persistent_queue_matches_block_per_poi(seq, targets):
    with mixed_driver(seq) as d:
        base = run_sequence(d, targets)
    with mixed_driver(seq, persistent=96) as d:
        pers = run_sequence(d, targets)
    for f, ((br, bs), (pr, ps)) in enumerate(zip(base, pers)):
        assert same_bytes(br, pr), f
        assert bs["iteration_sum"] == ps["iteration_sum"]


def test_fused_launch_matches_per_class(seq, targets):
    """One launch over all radius classes: parity gate against the per-class
    launches (other lane shapes, so other summation orders), and the
    persistent fused launch byte-identical to the one-block-per-POI one."""
    with mixed_driver(seq) as d:
        base = run_sequence(d, targets)
    with mixed_driver(seq) as d:
        d.set_fused(True)
        fused = run_sequence(d, targets)
    with mixed_driver(seq, persistent=128) as d:
        d.set_fused(True)
        fused_q = run_sequence(d, targets)
    for f, ((br, bs), (fr, fs), (qr, qs)) in enumerate(zip(base, fused, fused_q)):
        cmp = compare(br, fr)
        assert cmp["pass"], (f, cmp)
        assert same_bytes(fr, qr), f
        assert np.array_equal(fr["subset_radius"], br["subset_radius"])


def test_autotune_keeps_one_configuration(seq, targets):
    k = STEPS[-1]
    with mixed_driver(seq) as d:
        base, _ = d.process(img(targets[k]), policy=0)
    with mixed_driver(seq) as d:
        d.upload_target(img(targets[k]))
        rep = d.autotune()
        tuned, st = d.process(img(targets[k]), policy=0)
    fused_rows = [e for e in rep if e["radius"] == 0]
    assert fused_rows and {e["persistent"] > 0 for e in fused_rows} == {False, True}
    fused_chosen = any(e["chosen"] for e in fused_rows)
    for rad in synth.RADII_MIXED:
        rows = [e for e in rep if e["radius"] == rad]
        assert len(rows) >= 4
        # group sampling (2) exists for the one-wave shapes, which radii 10 and 15 have
        assert {e["sampling"] for e in rows} >= {0, 1} and {e["persistent"] > 0 for e in rows} == {False, True}
        assert (2 in {e["sampling"] for e in rows}) == any(e["waves"] == 1 for e in rows)
        assert all(e["persistent"] == 0 for e in rows if e["sampling"] == 2)
        assert sum(e["chosen"] for e in rows) == (0 if fused_chosen else 1)
        say("radius", rad, "best", min(rows, key=lambda e: e["ms"]), "of", len(rows))
    say("fused best", min(fused_rows, key=lambda e: e["ms"]), "chosen" if fused_chosen else "not chosen")
    assert sum(e["chosen"] for e in rep) == (1 if fused_chosen else len(synth.RADII_MIXED))
    cmp = compare(base, tuned)
    assert cmp["pass"], cmp


def test_frame_option_errors(seq, targets):
    """Error behaviour of the multi-frame options (status codes through the ABI)."""
    q = seq.pois[:200].copy()
    q["subset_radius"][5, 1] = 12.0  # not square
    with fa.FrameDriver(15) as d:
        d.set_self_adaptive(True)
        d.set_reference(img(seq.ref))
        d.set_pois(q)
        d.upload_target(img(targets[0]))
        with pytest.raises(NotImplementedError):
            d.run(0)
        q["subset_radius"][5] = 40.0  # beyond the device limit
        d.set_pois(q)
        with pytest.raises(NotImplementedError):
            d.run(0)
    with fa.FrameDriver(15) as d:
        d.set_reference(img(seq.ref))
        d.set_pois(seq.pois[:200].copy())
        with pytest.raises(fa.LogicError):
            d.autotune()            # no target uploaded
        with pytest.raises(fa.LogicError):
            d.download_redo()       # no frame run
        with pytest.raises(ValueError):
            d.set_persistent(-1)
        d.upload_target(img(targets[0]))
        d.set_fused(True)           # one class: the fused launch is the per-class one
        res, st = d.process(img(targets[0]))  # first frame: FFTCC, sets the bad-ratio baseline
        assert st["skip_fftcc_used"] == 0 and d.download_redo().all()
        res2, st2 = d.process(img(targets[8]))
        assert st2["skip_fftcc_used"] == 1
        assert d.download_redo().sum() == st2["poi_fallback_count"]


def test_streamed_uploads_match_serial_frames(seq, targets):
    """Target k+1 queued while frame k runs (two slots, background copy):
    byte-identical to uploading and running one frame at a time; a third
    queued target is refused; a 12-bit image above 4095 is refused by the run
    that takes it (the staging thread checks it as it copies), and the driver
    runs the next target."""
    with mixed_driver(seq) as d:
        serial = run_sequence(d, targets)
    with mixed_driver(seq, graph=True) as d:
        streamed = []
        d.upload_target(img(targets[STEPS[0]]))
        for i, k in enumerate(STEPS):
            if i + 1 < len(STEPS):
                d.upload_target(img(targets[STEPS[i + 1]]))
            st = d.run()
            streamed.append((d.download(), st))
        d.upload_target(img(targets[STEPS[0]]))
        d.upload_target(img(targets[STEPS[1]]))
        with pytest.raises(fa.LogicError):
            d.upload_target(img(targets[STEPS[2]]))
        bad = targets[STEPS[0]].copy()
        bad[3, 3] = 5000
        d.run()
        d.run()
        d.upload_target(img(bad))
        with pytest.raises(ValueError):
            d.run()
        d.upload_target(img(targets[STEPS[0]]))
        d.run()
    for f, ((a, _), (b, _)) in enumerate(zip(serial, streamed)):
        assert same_bytes(a, b), f


def test_graph_replay_with_persistent_queue_and_learned_seed(seq, targets):
    """Graph replay byte-identical to eager frames with the persistent work
    queue (its counters zeroed by a captured fill kernel) and with the
    learned seeder inside the graph."""
    with mixed_driver(seq, persistent=96) as d:
        eager = run_sequence(d, targets)
    with mixed_driver(seq, persistent=96, graph=True) as d:
        graph = run_sequence(d, targets)
    for f, ((er, _), (gr, _)) in enumerate(zip(eager, graph)):
        assert same_bytes(er, gr), f
    from fasticgn_amd import seeder as sd
    sc = synth.config0()
    outs = []
    for use_graph in (False, True):
        with sd.LearnedSeeder() as s, fa.FrameDriver(15) as d:
            d.set_reference(fa.Image2D(sc.ref))
            d.set_pois(sc.pois)
            d.set_seeder(s, 0.8)
            d.set_graph(use_graph)
            runs = [d.process(fa.Image2D(sc.tar), policy=fa.driver.LEARNED_SEED) for _ in range(3)]
        outs.append(runs)
    for (er, es), (gr, gs) in zip(*outs):
        assert same_bytes(er, gr)
        assert es["seed_fallback_count"] == gs["seed_fallback_count"]


def test_frame_large_strain_samples_target_without_table():
    """20 % strain with the gradients seeded: the warped neighbourhoods outgrow
    the LDS tile.  The frame driver builds no B-spline table unless a launch
    samples it, so those iterations read the f32 target in global memory; in
    table mode they read the table.  Both meet the parity gate against the
    CPU oracle's FFTCC + IC-GN from the same template."""
    from oracle import pyoracle as oracle
    sc = synth.affine_scene("strain", 384, 384, 3400, 7, t=(1.2, -0.7), grad=(0.2, 0.0, 0.0, 0.2),
                            radius=15, grid=16)
    q = sc.pois.copy()
    u, v, _, _, _, _ = sc.truth(q["x"], q["y"])
    q["deformation"][:, 0] = np.round(u)
    q["deformation"][:, 6] = np.round(v)
    q["deformation"][:, 1] = 0.2
    q["deformation"][:, 8] = 0.2
    o = q.copy()
    ref32, tar32 = sc.ref.astype(np.float32), sc.tar.astype(np.float32)
    oracle.fftcc(ref32, tar32, 15, o)
    oracle.icgn(ref32, tar32, 15, o)
    for mode in ("auto", "table"):
        with fa.FrameDriver(15) as d:
            d.set_reference(fa.Image2D(sc.ref))
            d.set_pois(q)
            d.set_sampling(mode)
            g, _ = d.process(fa.Image2D(sc.tar), policy=0)
        cmp = compare(o, g)
        say("large strain", mode, {k: cmp[k] for k in ("n", "gpu_valid", "status_mismatch", "p99_du", "max_du")})
        assert cmp["pass"], cmp
        assert cmp["gpu_valid"] > 0.5 * cmp["n"]


@pytest.mark.parametrize("mode", ["group", "pair"])
def test_frame_group_sampling_meets_the_gate(seq, targets, mode):
    """Group sampling (four neighbouring POIs per block sharing a coefficient
    tile) through the mixed-radius driver: radii 10 and 15 run one-wave shapes
    in group mode, radius 20 runs as 'auto'; pair sampling runs radius 15 two
    POIs per wave (the other radii as group sampling).  The frame meets the
    parity gate against the default sampling, and graph replay is
    byte-identical."""
    k = STEPS[2]
    with mixed_driver(seq) as d:
        base, _ = d.process(img(targets[k]), policy=0)
    with mixed_driver(seq) as d:
        d.set_sampling(mode)
        grp, st = d.process(img(targets[k]), policy=0)
    cmp = compare(base, grp)
    say(f"{mode} vs auto", {x: cmp[x] for x in ("n", "gpu_valid", "status_mismatch", "p99_du", "max_du")})
    assert cmp["pass"], cmp
    with mixed_driver(seq, graph=True) as d:
        d.set_sampling(mode)
        a, _ = d.process(img(targets[k]), policy=0)
        d.reset_sequence()
        b, _ = d.process(img(targets[k]), policy=0)
    assert same_bytes(a, b) and same_bytes(a, grp)
