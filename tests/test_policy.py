"""Pins of the frame-policy oracle (tests/policy.py), on the CPU.

1. The four predicates and the seeding of cli:493-624 on hand-built POIs, one
   case per branch (-3, NaN, zncc < 0, moved POI, the extrapolation rules).
2. The whole-frame fallback scenarios of smoke/run_frame_fallback_gate.py
   (tests/fallback_scenes.py) through the policy oracle: the gate's six
   assertions hold, and the two wrong merges the reference's gate was written
   against fail them (a fallback that overwrites the skip attempt fails
   `cascade`; a merge that never replaces a -3 fails `jump`), so the scenarios
   have the power to catch a wrong merge in the device driver too.
"""
import os

import numpy as np
import pytest

import tests.policy as pol
from fasticgn_amd.opencorr import poi_queue
from tests import fallback_scenes as fs

NAN = float("nan")
# the render seeds of the two cases (the occluder of `cascade` leaves 1-2 of
# 5,481 POIs on a false match with zncc >= 0.8 on one of eight seeds tried,
# 20261204; the others pass every assertion)
SEEDS = {"cascade": 20261211, "jump": 20261204}


def pois(zs, xy=None, u=None, v=None):
    q = poi_queue(xy if xy is not None else [(10 + i, 20) for i in range(len(zs))])
    q["result"][:, 2] = zs
    if u is not None:
        q["deformation"][:, 0] = u
    if v is not None:
        q["deformation"][:, 6] = v
    return q


# ------------------------------------------------------------ predicates --

def test_is_bad():
    """isBadPoiForFallback: -3 never counts; NaN, inf, -4, -5 and zncc < 0.8 do."""
    z = [-3.0, NAN, np.inf, -4.0, -5.0, 0.79, 0.8, 0.99]
    assert pol.is_bad(pois(z)).tolist() == [False, True, True, True, True, True, False, False]


def test_should_replace():
    """shouldReplacePoiWithFallback: a NaN or -3 fallback never replaces; a
    non-finite current is always replaced; else the fallback's zncc >= the
    current's."""
    cur = pois([0.5, 0.5, NAN, 0.5, 0.5, -5.0, 0.7])
    fb = pois([NAN, -3.0, -4.0, 0.5, 0.49, -4.0, 0.9])
    assert pol.should_replace(cur, fb).tolist() == [False, False, True, True, False, True, True]


def test_take_frame_fallback():
    """shouldTakeFrameFallbackResult: a -3 skip result only by a good scratch
    result; a good skip result never; a bad one under the repair rule."""
    skip = pois([-3.0, -3.0, -3.0, -3.0, 0.9, 0.5, 0.5, NAN, -4.0])
    scr = pois([0.85, 0.79, NAN, -3.0, 0.99, 0.6, 0.4, -4.0, -5.0])
    assert pol.take_frame_fallback(skip, scr).tolist() == [True, False, False, False, False, True, False,
                                                           True, False]


def test_is_candidate():
    """isPoiLevelFallbackCandidate: only bad POIs; a failed or negative result
    always; else unless the previous frame was bad at the same location."""
    cur = pois([0.9, NAN, -4.0, 0.5, 0.5, 0.5, 0.5, 0.5])
    prev = pois([0.9, 0.9, 0.9, 0.9, 0.5, NAN, 0.9, 0.5])
    prev["x"][6] += 1  # moved POI
    bad = pol.is_bad(cur)
    assert pol.is_candidate(cur, prev, bad).tolist() == [False, True, True, True, False, True, True, False]
    assert pol.is_candidate(cur, None, bad).tolist() == [False, True, True, True, True, True, True, True]


def test_seed_from_previous_copies_only_valid_results():
    t = pois([0.0] * 4)
    t["deformation"][:, 0] = 7.0
    t["result"][:, 3] = 9.0
    prev = pois([NAN, -0.1, 0.0, 0.95], u=[1.0, 2.0, 3.0, 4.0], v=[-1.0, -2.0, -3.0, -4.0])
    prev["deformation"][:, 1] = 0.25
    prev["result"][:, 3] = 4.0
    q, count = pol.seed_from_previous(t, prev)
    assert count == 0
    assert q["deformation"][:, 0].tolist() == [7.0, 7.0, 3.0, 4.0]
    assert q["deformation"][:, 1].tolist() == [0.0, 0.0, 0.25, 0.25]
    assert q["result"][:, 3].tolist() == [9.0, 9.0, 0.0, 0.0]  # result cleared where seeded


def test_temporal_extrapolation_rules():
    """seedPoiQueueFromPrevious with extrapolation (cli:597-615): u + (u - u_prev2)
    when both frames are good at the same location, the step is at least
    0.05 px and each component at most max_delta."""
    xy = [(10 + i, 20) for i in range(8)]
    t = pois([0.0] * 8, xy)
    prev = pois([0.9, 0.9, 0.9, 0.9, 0.79, 0.9, 0.9, 0.9], xy,
                u=[2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, NAN], v=[1.0] * 8)
    prev2 = pois([0.9, 0.9, 0.9, 0.79, 0.9, 0.9, 0.9, 0.9], xy,
                 u=[1.5, 1.97, -14.0, 1.5, 1.5, 1.5, 1.5, 1.5], v=[1.0] * 8)
    prev2["x"][5] += 1      # moved between the two frames
    q, count = pol.seed_from_previous(t, prev, prev2, True, 15.0)
    # 0: extrapolated; 1: step 0.03 < 0.05; 2: |du| 16 > 15; 3: prev2 zncc 0.79;
    # 4: prev zncc 0.79 (seeded, not extrapolated); 5: moved; 6: extrapolated; 7: NaN u
    assert q["deformation"][:, 0].tolist()[:7] == [2.5, 2.0, 2.0, 2.0, 2.0, 2.0, 2.5]
    assert count == 2
    off, _ = pol.seed_from_previous(t, prev, prev2, False, 15.0)
    assert off["deformation"][0, 0] == 2.0
    # the v step alone counts, and the step test is >= 0.05 in float32
    prev2b = pois([0.9] * 8, xy, u=[2.0] * 8, v=[0.95] * 8)
    q, count = pol.seed_from_previous(t, prev, prev2b, True, 15.0)
    step = np.float32(np.float32(1.0) - np.float32(0.95))
    assert (step * step >= pol.MIN_STEP * pol.MIN_STEP) == (q["deformation"][0, 6] != 1.0)


# -------------------------------------------------------------- scenarios --

def run_case(oracle, case, seed, merge=None):
    ref, tars = fs.sequence(case, seed)
    po = pol.PolicyOracle(oracle, ref, fs.template(case), case.radius)
    frames, stats = [], []
    for tar in tars:
        q, st = po.frame(tar)
        frames.append(q)
        stats.append(st)
    return frames, stats


@pytest.mark.parametrize("name", ["cascade", "jump"])
def test_fallback_scenario_assertions(oracle, name):
    case = fs.CASES[name]
    frames, stats = run_case(oracle, case, SEEDS[name])
    errors, report = fs.check_case(case, frames, stats, [s.get("scratch") for s in stats])
    for row in report:
        print(row)
    assert not errors, errors
    # assertion 4: the frames do not depend on the oracle's thread count
    old = os.environ.get("ORACLE_THREADS")
    os.environ["ORACLE_THREADS"] = "1"
    try:
        frames1, _ = run_case(oracle, case, SEEDS[name])
    finally:
        if old is None:
            del os.environ["ORACLE_THREADS"]
        else:
            os.environ["ORACLE_THREADS"] = old
    assert all(a.tobytes() == b.tobytes() for a, b in zip(frames, frames1))


@pytest.mark.parametrize("name,mutant,expect", [
    ("cascade", "overwrite", "raised the bad ratio"),
    ("jump", "never_minus3", "valid coverage"),
])
def test_fallback_scenarios_catch_a_wrong_merge(oracle, monkeypatch, name, mutant, expect):
    orig = pol.take_frame_fallback
    mutants = {
        "overwrite": lambda skip, scratch: np.ones(len(skip), bool),
        "never_minus3": lambda skip, scratch: np.where(skip["result"][:, 2] == -3.0, False, orig(skip, scratch)),
    }
    monkeypatch.setattr(pol, "take_frame_fallback", mutants[mutant])
    case = fs.CASES[name]
    frames, stats = run_case(oracle, case, SEEDS[name])
    errors, _ = fs.check_case(case, frames, stats, [s.get("scratch") for s in stats])
    assert any(expect in e for e in errors), errors
