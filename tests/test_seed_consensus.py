"""Neighbourhood consensus of the learned seeds.

CPU: the restatement (oracle/seeder_ref.consensus) pinned by closed forms: a
linear field carried exactly over any neighbourhood, including the ragged
ones at the edge of the POI set; the plain window mean of a symmetric
neighbourhood with zero gradients (the +-radius window, not a larger one);
a step of the field left out by the tau test; unseeded POIs and non-finite
positions or predictions neither changed nor used.

GPU: the device consensus (fi_seeder_seed_consensus) bit-identical to the
restatement applied to the device's raw seeds; the frame driver's learned
path with consensus meets the parity gate against the oracle run from the
same seeds, with the same iteration counts, and needs fewer IC-GN iterations
than the raw seeds on a smooth field."""
import numpy as np
import pytest

import fasticgn_amd as fa
from fasticgn_amd import synth
from oracle import seeder_ref as S


def grid(x0, x1, y0, y1, step):
    x, y = np.meshgrid(np.arange(x0, x1, step, dtype=np.float64), np.arange(y0, y1, step, dtype=np.float64))
    return x.ravel().astype(np.float32), y.ravel().astype(np.float32)


def linear_seeds(x, y, a):
    """u = a0 + a1 x + a2 y, v = a3 + a4 x + a5 y with their exact gradients."""
    s = np.zeros((len(x), 6), np.float64)
    s[:, 0] = a[0] + a[1] * x + a[2] * y
    s[:, 1] = a[3] + a[4] * x + a[5] * y
    s[:, 2], s[:, 3], s[:, 4], s[:, 5] = a[1], a[2], a[4], a[5]
    return s.astype(np.float32)


# ------------------------------------------------------------------ CPU --

def test_linear_field_is_reproduced_everywhere():
    """Every neighbour's u, v carried to the POI equals the POI's own, and the
    centroid correction with the mean gradients is exact, so the field comes
    back unchanged at interior, edge and corner POIs alike (FP32 rounding)."""
    x, y = grid(20, 220, 30, 200, 5)
    a = (1.25, 0.013, -0.021, -3.5, 0.034, 0.0045)
    s = linear_seeds(x, y, a)
    o = S.consensus(x, y, s, np.ones(len(x), bool), 15.0, 0.5)
    np.testing.assert_allclose(o[:, :2], s[:, :2], atol=2e-6)
    np.testing.assert_allclose(o[:, 2:], s[:, 2:], atol=1e-7)
    # irregular positions too
    rng = np.random.default_rng(4)
    xr = rng.uniform(0, 300, 700).astype(np.float32)
    yr = rng.uniform(0, 300, 700).astype(np.float32)
    sr = linear_seeds(xr, yr, a)
    orr = S.consensus(xr, yr, sr, np.ones(700, bool), 15.0, 0.5)
    np.testing.assert_allclose(orr[:, :2], sr[:, :2], atol=3e-6)


def test_symmetric_window_is_the_plain_mean():
    """Zero gradients, random u, v: an interior POI of a 5 px grid with radius
    15 takes the mean over exactly its 7 x 7 window (a radius of 14.9 would
    take 5 x 5)."""
    x, y = grid(0, 200, 0, 200, 5)
    nx = 40
    rng = np.random.default_rng(7)
    s = np.zeros((len(x), 6), np.float32)
    s[:, 0] = rng.normal(0, 0.05, len(x))
    s[:, 1] = rng.normal(0, 0.05, len(x))
    o = S.consensus(x, y, s, np.ones(len(x), bool), 15.0, 0.5)
    u = s[:, 0].reshape(40, 40)
    for gy, gx in ((10, 10), (20, 31), (5, 3)):
        i = gy * nx + gx
        win = u[gy - 3:gy + 4, gx - 3:gx + 4]
        assert abs(o[i, 0] - win.mean()) < 1e-6
    o2 = S.consensus(x, y, s, np.ones(len(x), bool), 14.9, 0.5)
    i = 10 * nx + 10
    assert abs(o2[i, 0] - u[8:13, 8:13].mean()) < 1e-6


def test_step_is_not_averaged_across():
    """A 1.5 px step of u along x = 100: every POI's consensus equals its own
    side's constant (neighbours on the other side disagree by more than tau);
    with tau above the step the POIs next to it are pulled toward it."""
    x, y = grid(0, 200, 0, 100, 5)
    s = np.zeros((len(x), 6), np.float32)
    s[:, 0] = np.where(x < 100, 2.0, 3.5)
    s[:, 1] = -1.0
    o = S.consensus(x, y, s, np.ones(len(x), bool), 15.0, 0.5)
    np.testing.assert_allclose(o[:, 0], s[:, 0], atol=1e-6)  # a mean of equal values, up to FP32 rounding
    np.testing.assert_allclose(o[:, 1], s[:, 1], atol=1e-6)
    wide = S.consensus(x, y, s, np.ones(len(x), bool), 15.0, 2.0)
    near = np.abs(x - 100) < 10
    assert np.all(np.abs(wide[near, 0] - s[near, 0]) > 0.1)


def test_unseeded_and_non_finite_are_left_out():
    x, y = grid(0, 100, 0, 100, 5)
    n = len(x)
    s = linear_seeds(x, y, (0.5, 0.0, 0.0, 0.25, 0.0, 0.0))
    ok = np.ones(n, bool)
    ok[::7] = False
    s[::7, 0] = 40.0      # unseeded: must not pull its neighbours
    s[3, 1] = np.nan      # a NaN prediction: left as it is, not used
    x = x.copy()
    x[5] = np.nan         # a POI without a finite position: unchanged
    o = S.consensus(x, y, s, ok, 15.0, 0.5)
    assert np.all(o[~ok] == s[~ok])
    assert np.isnan(o[3, 1]) and o[3, 0] == s[3, 0]
    assert np.all(o[5] == s[5])
    good = ok.copy()
    good[[3, 5]] = False
    np.testing.assert_allclose(o[good, 0], 0.5, atol=1e-6)


@pytest.mark.parametrize("mode", [1, 2])
def test_sampled_seeding_linear_field(mode):
    """The network seeds a checkerboard (1) or a quarter (2) of the cells; the
    other POIs take their neighbours' consensus, carried to them: a linear
    field comes back everywhere and every POI is seeded.  The sample is the
    cell pattern (cells of side radius / 3 = one grid step here)."""
    x, y = grid(20, 220, 30, 200, 5)
    a = (1.25, 0.013, -0.021, -3.5, 0.034, 0.0045)
    s = linear_seeds(x, y, a)
    cell, ncx, *_ = S.seed_cells(x, y, 15.0)
    picked = S.seed_sample(cell, ncx, mode)
    assert abs(picked.mean() - (0.5 if mode == 1 else 0.25)) < 0.03
    seeds = np.where(picked[:, None], s, np.float32(0))  # the others have no prediction
    ok = picked.copy()
    o, okk = S.consensus(x, y, seeds, ok, 15.0, 0.5, sample_mode=mode, return_ok=True)
    assert okk.all()
    np.testing.assert_allclose(o[:, :2], s[:, :2], atol=3e-6)
    np.testing.assert_allclose(o[:, 2:], s[:, 2:], atol=1e-7)


def test_sampled_seeding_at_a_step():
    """A POI outside the sample measures tau from the mean of its neighbours'
    predictions: next to a 1.5 px step it takes its majority side's value
    exactly, or, with a neighbourhood split near half and half, no seed at all
    (FFTCC fallback) - never a blend of the two sides."""
    x, y = grid(0, 200, 0, 100, 5)
    s = np.zeros((len(x), 6), np.float32)
    s[:, 0] = np.where(x < 100, 2.0, 3.5)
    s[:, 1] = -1.0
    cell, ncx, *_ = S.seed_cells(x, y, 15.0)
    picked = S.seed_sample(cell, ncx, 1)
    o, okk = S.consensus(x, y, s, picked.copy(), 15.0, 0.5, sample_mode=1, return_ok=True)
    other = ~picked
    side = np.where(x < 100, 2.0, 3.5)
    assert np.all(np.abs(o[other & okk, 0] - side[other & okk]) < 1e-6)
    assert (other & ~okk).any() and (other & okk & (np.abs(x - 100) <= 15)).any()
    assert np.all(okk[np.abs(x - 100) > 15])


# ------------------------------------------------------------------ GPU --

@pytest.mark.gpu
@pytest.mark.parametrize("sample", ["all", "checker", "quarter"])
def test_device_consensus_matches_restatement(sample):
    """fi_seeder_seed_consensus on a config-1 slice against the restatement
    applied to the raw device seeds: identical bits and flags, with the network
    on every POI or on a sample of them."""
    from fasticgn_amd import seeder as sd
    sc = synth.config1(size=512)
    ref32, tar32 = sc.ref.astype(np.float32), sc.tar.astype(np.float32)
    q = sc.pois.copy()
    q["x"][7] = np.nan  # a POI without a finite position is left alone
    with sd.LearnedSeeder() as s:
        raw = q.copy()
        ok = s.seed(ref32, tar32, raw)
        dev = q.copy()
        ok2 = s.seed(ref32, tar32, dev, consensus=(15.0, 0.5), sample=sample)
    cols = [0, 6, 1, 2, 7, 8]
    d = raw["deformation"]
    seeds = np.stack([d[:, c] for c in cols], 1)
    want, want_ok = S.consensus(raw["x"], raw["y"], seeds, ok, 15.0, 0.5,
                                sample_mode=fa._native.SEED_SAMPLES[sample], return_ok=True)
    got = np.stack([dev["deformation"][:, c] for c in cols], 1)
    assert np.array_equal(ok2, want_ok)
    m = want_ok  # a POI left unseeded keeps whatever it held
    assert got[m].tobytes() == want[m].tobytes()
    assert np.abs(got - seeds)[ok & m].max() > 0  # the consensus changed the seeds


@pytest.mark.gpu
def test_frame_consensus_parity_and_iterations():
    """The learned frame with consensus, the network on a checkerboard of the
    POIs, against the oracle run from the same seeds on a POI sample: the
    parity gate and identical iteration counts up to summation-order noise;
    fewer IC-GN iterations than the raw seeds, with or without the sample."""
    from fasticgn_amd import seeder as sd
    from oracle import pyoracle
    from tests.parity import compare
    from tests.policy import PolicyOracle
    sc = synth.config1(size=1024)
    stats = {}
    with sd.LearnedSeeder() as s:
        for on, sample in ((False, "all"), (True, "all"), (True, "checker")):
            with fa.FrameDriver(sc.radius) as d:
                d.set_reference(fa.Image2D(sc.ref))
                d.set_pois(sc.pois)
                d.set_seeder(s, 0.8)
                d.set_seed_consensus(on, 0.5, sample)
                d.upload_target(fa.Image2D(sc.tar))
                stats[on, sample] = d.run(fa.driver.LEARNED_SEED)
                if sample == "checker":
                    res = d.download()
        print("iterations per POI: raw", stats[False, "all"]["iteration_mean"], "consensus",
              stats[True, "all"]["iteration_mean"], "checkerboard", stats[True, "checker"]["iteration_mean"],
              flush=True)
        assert stats[True, "all"]["iteration_mean"] < stats[False, "all"]["iteration_mean"] - 0.15
        assert stats[True, "checker"]["iteration_mean"] < stats[False, "all"]["iteration_mean"] - 0.15
        ref32, tar32 = sc.ref.astype(np.float32), sc.tar.astype(np.float32)
        seeds = sc.pois.copy()
        ok = s.seed(ref32, tar32, seeds, consensus=(float(sc.radius), 0.5), sample="checker")
    pick = np.arange(0, len(sc.pois), 9)
    po = PolicyOracle(pyoracle, sc.ref, sc.pois[pick], sc.radius)
    o, _ = po.learned_frame(sc.tar, seeds[pick], ok[pick], 0.8)
    cmp = compare(o, res[pick])
    print({k: cmp[k] for k in ("n", "status_mismatch", "p99_du", "max_du", "iteration_mismatch_frac")}, flush=True)
    assert cmp["pass"], cmp
    assert cmp["iteration_mismatch_frac"] < 0.01
