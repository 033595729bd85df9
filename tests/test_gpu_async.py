"""Frames launched behind a running frame (fi_frame_run_async / fi_frame_wait).

A frame that needs nothing from the one before (FFTCC seeds from the
template, or learned seeds, replayed as graphs) is queued on the device
while the previous frame still runs, on the other set of per-frame buffers.
Frame for frame its records, its statistics and the final results must be
byte-identical to fi_frame_run's, with the targets changing every frame (a
target-slot or buffer-set mix-up shows as another frame's data), and with
the learned frame's ZNCC threshold raised so that most POIs are repair
candidates: the candidates past the device's repair cap are then repaired
by the host when the frame completes, while the next frame already runs.
"""
import numpy as np
import pytest

import fasticgn_amd as fa
from fasticgn_amd import synth

pytestmark = pytest.mark.gpu

LEARNED = fa.driver.LEARNED_SEED


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def scene():
    sc = synth.config0()
    tars = [sc.tar, np.roll(sc.tar, 1, axis=1), np.roll(sc.tar, -1, axis=0), sc.tar, np.roll(sc.tar, 2, axis=1)]
    return sc, tars


def stats_key(st):
    """The statistics but the timings; NaN (a ratio with no baseline yet) as
    a string, so that equal statistics compare equal."""
    def val(v):
        return "nan" if isinstance(v, float) and v != v else v
    return tuple((k, val(st[k])) for k in sorted(st) if not k.endswith("_ms"))


def driver(sc, seeder, threshold):
    d = fa.FrameDriver(15)
    d.set_reference(fa.Image2D(sc.ref))
    d.set_pois(sc.pois)
    d.set_graph(True)
    d.set_records(True)
    if seeder is not None:
        d.set_seeder(seeder, threshold)
    return d


def serial(sc, tars, policy, seeder=None, threshold=0.8):
    recs, stats, res = [], [], []
    with driver(sc, seeder, threshold) as d:
        for t in tars:
            img = fa.Image2D(t)
            d.upload_target(img)
            stats.append(stats_key(d.run(policy)))
            recs.append(d.records(0).tobytes())
            res.append(d.download().tobytes())
    return recs, stats, res


def pipelined(sc, tars, policy, seeder=None, threshold=0.8, peek=None):
    """run_async(k), wait(k - 1) with its records (age 1), upload(k + 1);
    at frame `peek` the results are downloaded right after its launch (the
    download completes the frames in flight first)."""
    imgs = [fa.Image2D(t) for t in tars]
    recs, stats, peeked = [], [], None
    with driver(sc, seeder, threshold) as d:
        d.upload_target(imgs[0])
        for k in range(len(tars)):
            d.run_async(policy)
            if k == peek:
                peeked = d.download().tobytes()
            if k > 0:
                stats.append(stats_key(d.wait()))
                recs.append(d.records(1).tobytes())
            if k + 1 < len(tars):
                d.upload_target(imgs[k + 1])
        stats.append(stats_key(d.wait()))
        recs.append(d.records(0).tobytes())
        last = d.download().tobytes()
    return recs, stats, last, peeked


@pytest.mark.parametrize("mode", ["fftcc", "learned", "learned-host-repair"])
def test_pipelined_frames_match_serial(scene, mode):
    from fasticgn_amd import seeder as sd
    sc, tars = scene
    policy = 0 if mode == "fftcc" else LEARNED
    threshold = 0.9999 if mode == "learned-host-repair" else 0.8
    say("mode", mode, "scene", sc.ref.shape, len(sc.pois), "POIs,", len(tars), "targets")
    with sd.LearnedSeeder() as s:
        seeder = None if mode == "fftcc" else s
        recs, stats, res = serial(sc, tars, policy, seeder, threshold)
        precs, pstats, last, peeked = pipelined(sc, tars, policy, seeder, threshold, peek=2)
    for k in range(len(tars)):
        assert precs[k] == recs[k], f"frame {k}: records differ"
        diff = [(a, b) for a, b in zip(pstats[k], stats[k]) if a != b]
        assert not diff, f"frame {k}: statistics differ (pipelined, serial): {diff}"
    assert last == res[-1]
    assert peeked == res[2]
    st = dict(stats[0])
    assert st["valid_count"] > 0.9 * len(sc.pois)
    if mode == "learned-host-repair":
        # most POIs are candidates: far past the device's repair cap (1024 here)
        assert st["seed_fallback_count"] > 1024, st
    # frames that differ in their targets give different records
    assert recs[0] != recs[1]


def test_async_call_order_errors(scene):
    sc, tars = scene
    with driver(sc, None, 0.8) as d:
        d.upload_target(fa.Image2D(tars[0]))
        with pytest.raises(Exception, match="fi_frame_run_async"):
            d.wait()
        d.run_async(0)
        with pytest.raises(Exception, match="fi_frame_wait"):
            d.run(0)
        st = d.wait()
        assert st["frame_index"] == 0
        # a plain run again once every launched frame has been waited for
        assert d.run(0)["frame_index"] == 1
