"""Parity of the gfx950 path (through the C ABI) with the CPU oracle.

Bars:
  FFTCC  integer peaks, zncc and every other POI field bit-identical to the
         oracle (both follow one DFT specification).
  ICGN   validity flags (valid / -3 / -4 / -5) identical; u, v and the
         displacement gradients ux, uy, vx, vy within the reference's parity
         gate, per frame median <= 1e-3, nearest-rank p99 <= 1e-2, max <= 0.5
         (smoke/check_cpu_gpu_parity.py:46-98, tests/parity.py); iteration
         counts and convergence agreement reported, and bounded on config 1.
"""
import numpy as np
import pytest

import fasticgn_amd as fa
from fasticgn_amd import synth
from fasticgn_amd.opencorr import poi_queue
from tests.parity import compare, finite_valid

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c0():
    return synth.config0()


def fields_equal(a, b):
    return a.tobytes() == b.tobytes()


# ------------------------------------------------------------------ FFTCC --

def test_fftcc_lcg_known_answer(oracle):
    ref, tar, q = synth.lcg_pair()
    say("fftcc KAT r=7 192x128", len(q), "POIs")
    R, T = fa.Image2D(ref), fa.Image2D(tar)
    with fa.FFTCC2D(7, 7) as f:
        f.setImages(R, T)
        f.prepareRef(q)
        assert f.hasPreparedReference(R)
        g = q.copy()
        f.compute(g)
    assert np.all(g["deformation"][:, 0] == 2) and np.all(g["deformation"][:, 6] == -1)
    assert np.all(g["result"][:, 2] > 0.5)
    o = q.copy()
    oracle.PreparedRef(ref, 7, q).compute(tar, o)
    assert fields_equal(g, o)


@pytest.mark.parametrize("radius", [7, 10, 15, 20])
def test_fftcc_bit_exact_cached_and_uncached(oracle, c0, radius):
    say("fftcc config0 radius", radius)
    ref, tar = c0.ref, c0.tar
    q = fa.grid_queue(512, 512, radius, 9)
    R, T = fa.Image2D(ref), fa.Image2D(tar)
    with fa.FFTCC2D(radius, radius) as f:
        f.setImages(R, T)
        f.prepareRef(q)
        g_cached = q.copy()
        f.compute(g_cached)
        moved = q.copy()
        moved["x"] += 1.0  # layout differs -> uncached path
        moved["deformation"][:, 0] = 1.0
        g_unc = moved.copy()
        f.compute(g_unc)
    o_cached = q.copy()
    oracle.PreparedRef(ref.astype(np.float32), radius, q).compute(tar.astype(np.float32), o_cached)
    o_unc = moved.copy()
    oracle.fftcc(ref.astype(np.float32), tar.astype(np.float32), radius, o_unc)
    assert fields_equal(g_cached, o_cached)
    assert fields_equal(g_unc, o_unc)
    assert (g_cached["result"][:, 2] > 0).mean() > 0.9


def test_fftcc_indexed_and_errors(oracle, c0):
    q = fa.grid_queue(512, 512, 15, 20)
    R, T = fa.Image2D(c0.ref), fa.Image2D(c0.tar)
    with fa.FFTCC2D(15, 15) as f:
        f.setImages(R, T)
        with pytest.raises(fa.LogicError):
            f.compute(q[:3].copy(), [0, 1, 2])  # no layout yet
        f.prepareRef(q)
        idx = np.array([5, 17, 3, 40])
        sub = q[idx].copy()
        f.compute(sub, idx)
        with pytest.raises(ValueError):
            f.compute(q[[1, 2]].copy(), [2, 1])  # coordinates do not match
        with pytest.raises(IndexError):
            f.compute(q[:1].copy(), [len(q)])
        with pytest.raises(ValueError):
            f.compute(q[:2].copy(), [0])
        small = fa.Image2D(np.zeros((64, 96), np.uint8))
        f.setTarImage(small)
        before = q.copy()
        with pytest.raises(ValueError, match="512x512.*96x64"):
            f.compute(before)
        assert fields_equal(before, q)
    o = q[idx].copy()
    oracle.PreparedRef(c0.ref.astype(np.float32), 15, q).compute(c0.tar.astype(np.float32), o, idx)
    assert fields_equal(sub, o)


def test_fftcc_empty_queue(c0):
    with fa.FFTCC2D(15, 15) as f:
        f.setImages(fa.Image2D(c0.ref), fa.Image2D(c0.tar))
        q = poi_queue(np.zeros((0, 2)))
        f.compute(q)
        f.prepareRef(q)
        f.compute(q)


# ------------------------------------------------------------------- ICGN --

def run_icgn_gpu(ref, tar, radius, q, conv=0.001, stop=10.0, self_adaptive=False, sampling="auto"):
    g = q.copy()
    with fa.ICGN2D1(radius, radius, conv, stop) as ic:
        ic.setSampling(sampling)
        if self_adaptive:
            ic.setSelfAdaptive(True)
        ic.setImages(fa.Image2D(ref), fa.Image2D(tar))
        ic.prepare()
        ic.compute(g)
    return g


@pytest.mark.parametrize("grid", [5, 10])
def test_icgn_group_sampling_parity(oracle, grid):
    """Group sampling: four consecutive POIs per block share one LDS tile of
    x-polynomial coefficients.  On a 5 px grid a block's neighbourhoods fit one
    tile; on a 10 px grid they do not and every sample reads the target in
    device memory (the table here).  Both meet the parity gate against the CPU
    oracle, validity flags identical."""
    sc = synth.affine_scene("group", 384, 384, 3400, 13, t=(1.6, -0.9), grad=(2e-3, 0.0, 0.0, 2e-3),
                            radius=15, grid=grid)
    q = sc.pois.copy()
    u, v, *_ = sc.truth(q["x"], q["y"])
    q["deformation"][:, 0] = np.round(u)
    q["deformation"][:, 6] = np.round(v)
    g = run_icgn_gpu(sc.ref, sc.tar, 15, q, sampling="group")
    o = q.copy()
    oracle.icgn(sc.ref.astype(np.float32), sc.tar.astype(np.float32), 15, o)
    cmp = compare(o, g)
    report(f"icgn group sampling grid {grid}", cmp)
    assert cmp["pass"], cmp
    assert finite_valid(g).mean() > 0.9


@pytest.mark.parametrize("grid,odd", [(5, True), (10, False)])
def test_icgn_pair_sampling_parity(oracle, grid, odd):
    """Two POIs per wave (lanes 0-31 and 32-63), eight neighbouring POIs per
    block sharing one coefficient tile.  On a 5 px grid a block's
    neighbourhoods fit the tile except where its eight POIs wrap to the next
    grid row; on a 10 px grid none fit and every sample reads the target in
    device memory.  An odd count leaves the last block's last half without a
    POI.  Parity gate against the CPU oracle, validity flags identical."""
    sc = synth.affine_scene("pair", 384, 384, 3400, 17, t=(1.4, -1.1), grad=(2e-3, 1e-3, -1e-3, 2e-3),
                            radius=15, grid=grid)
    n = len(sc.pois)
    q = sc.pois[:n if n % 2 == int(odd) else n - 1].copy()
    assert len(q) % 2 == int(odd)
    u, v, *_ = sc.truth(q["x"], q["y"])
    q["deformation"][:, 0] = np.round(u)
    q["deformation"][:, 6] = np.round(v)
    g = run_icgn_gpu(sc.ref, sc.tar, 15, q, sampling="pair")
    o = q.copy()
    oracle.icgn(sc.ref.astype(np.float32), sc.tar.astype(np.float32), 15, o)
    cmp = compare(o, g)
    report(f"icgn pair sampling grid {grid}", cmp)
    assert cmp["pass"], cmp
    assert finite_valid(g).mean() > 0.9
    # the same POIs through the shared-tile kernel: same statuses, the
    # displacements within summation-order noise
    grp = run_icgn_gpu(sc.ref, sc.tar, 15, q, sampling="group")
    both = compare(grp, g)
    assert both["status_mismatch"] == 0 and max(both["max_du"], both["max_dv"]) < 1e-3, both
    assert fields_equal(g, run_icgn_gpu(sc.ref, sc.tar, 15, q, sampling="pair"))


def test_icgn_center_offset_parity(oracle):
    """compute(queue, center_offset_queue) (src/oc_icgn.cpp:399-625) on the
    device against the CPU oracle's, offsets U[-3, 3] px on an affine field with
    gradients up to 0.05: the parity gate, u, v equal to the displacement at the
    offset centre, and zero offsets byte-identical to compute(queue)."""
    sc = synth.affine_scene("offset", 384, 384, 3400, 11, t=(1.3, -0.8), grad=(0.04, 0.02, -0.03, 0.05),
                            radius=15, grid=24)
    q = sc.pois.copy()
    u, v, ux, uy, vx, vy = sc.truth(q["x"], q["y"])
    d = q["deformation"]
    d[:, 0], d[:, 6] = np.round(u), np.round(v)
    d[:, 1], d[:, 2], d[:, 7], d[:, 8] = ux, uy, vx, vy
    off = np.random.default_rng(5).uniform(-3, 3, (len(q), 2)).astype(np.float32)
    with fa.ICGN2D1(15, 15, 0.001, 10) as ic:
        ic.setImages(fa.Image2D(sc.ref), fa.Image2D(sc.tar))
        ic.prepare()
        plain, zero, g = q.copy(), q.copy(), q.copy()
        ic.compute(plain)
        ic.compute(zero, np.zeros((len(q), 2), np.float32))
        ic.compute(g, off)
        again = q.copy()
        ic.compute(again)  # the prepared records are rebuilt after an offset call
    assert plain.tobytes() == zero.tobytes()
    assert plain.tobytes() == again.tobytes()
    o = q.copy()
    oracle.icgn(sc.ref.astype(np.float32), sc.tar.astype(np.float32), 15, o, center_offsets=off)
    cmp = compare(o, g)
    report("icgn center offset", cmp)
    assert cmp["pass"], cmp
    ok = finite_valid(g)
    ut, vt, *_ = sc.truth(g["x"] + off[:, 0], g["y"] + off[:, 1])
    err = np.hypot(g["deformation"][ok, 0] - ut[ok], g["deformation"][ok, 6] - vt[ok])
    say("center offset vs truth at the offset centre: p99", float(np.percentile(err, 99)))
    assert ok.mean() > 0.8 and np.percentile(err, 99) < 0.03
    with fa.ICGN2D1(15, 15, 0.001, 10) as ic:
        ic.setImages(fa.Image2D(sc.ref), fa.Image2D(sc.tar))
        ic.prepare()
        with pytest.raises(ValueError):
            ic.compute(q.copy(), off[:-1])


def report(tag, cmp):
    keys = ["n", "cpu_valid", "gpu_valid", "status_mismatch", "median_du", "p99_du", "max_du",
            "median_dv", "p99_dv", "max_dv", "p99_dux", "p99_duy", "p99_dvx", "p99_dvy", "max_dux", "max_dvy",
            "iteration_mismatch", "iteration_mismatch_frac", "max_diteration", "max_dconvergence"]
    say(tag, {k: cmp.get(k) for k in keys})


def test_icgn_config0_parity(oracle, c0):
    say("icgn config0 r=15", len(c0.pois), "POIs")
    ref, tar = c0.ref, c0.tar
    seeds = c0.pois.copy()
    oracle.fftcc(ref.astype(np.float32), tar.astype(np.float32), 15, seeds)
    g = run_icgn_gpu(ref, tar, 15, seeds)
    o = seeds.copy()
    oracle.icgn(ref.astype(np.float32), tar.astype(np.float32), 15, o)
    cmp = compare(o, g)
    report("icgn config0 vs FP32 oracle", cmp)
    assert cmp["pass"], cmp
    o64 = seeds.copy()
    oracle.icgn(ref.astype(np.float32), tar.astype(np.float32), 15, o64, f64=True)
    c64 = compare(o64, g)
    report("icgn config0 vs FP64 oracle", c64)
    assert c64["pass"]
    ok = finite_valid(g)
    u, v, *_ = c0.truth(g["x"], g["y"])
    err = np.hypot(g["deformation"][ok, 0] - u[ok], g["deformation"][ok, 6] - v[ok])
    say("icgn config0 vs truth rms", float(np.sqrt(np.mean(err ** 2))))
    assert np.sqrt(np.mean(err ** 2)) < 0.03
    # untouched fields pass through
    np.testing.assert_array_equal(g["strain"], seeds["strain"])
    np.testing.assert_array_equal(g["deformation"][:, [3, 4, 5, 9, 10, 11]],
                                  seeds["deformation"][:, [3, 4, 5, 9, 10, 11]])


@pytest.mark.parametrize("radius", [7, 10, 20])
def test_icgn_other_radii_parity(oracle, c0, radius):
    say("icgn config0 radius", radius)
    ref, tar = c0.ref, c0.tar
    q = fa.grid_queue(512, 512, radius, 13)
    oracle.fftcc(ref.astype(np.float32), tar.astype(np.float32), radius, q)
    g = run_icgn_gpu(ref, tar, radius, q)
    o = q.copy()
    oracle.icgn(ref.astype(np.float32), tar.astype(np.float32), radius, o)
    cmp = compare(o, g)
    report(f"icgn r={radius}", cmp)
    assert cmp["pass"], cmp


def test_icgn_rejects_images_beyond_the_offset_range():
    """The IC-GN kernels address the f32 images with 32-bit offsets: a
    reference above 2 GiB as f32 (23171^2 pixels) is refused when it is set,
    before any pixel is read or uploaded; 23170^2 is the largest square."""
    assert 23170 * 23170 * 4 <= 2**31 - 1 < 23171 * 23171 * 4
    big = fa.Image2D(np.zeros((23171, 23171), np.uint8))  # calloc: never touched
    with fa.ICGN2D1(15, 15, 0.001, 10) as ic:
        with pytest.raises(NotImplementedError, match="2 GiB"):
            ic.setRefImage(big)
    with fa.FrameDriver(15) as d:
        with pytest.raises(NotImplementedError, match="2 GiB"):
            d.set_reference(big)


@pytest.mark.parametrize("sampling", ["auto", "pair"])
def test_icgn_edge_cases(oracle, c0, sampling):
    """OOB band, zero texture (-5), entry gate: statuses identical, and the
    fields the reference leaves alone are left alone.  In pair sampling the two
    POIs of a wave meet different fates (one rejected at the gate or leaving the
    image while the other iterates on)."""
    ref = c0.ref.copy()
    ref[100:200, 100:200] = 128
    tar = np.roll(ref, (-2, 3), axis=(0, 1))
    pts = [(x, y) for y in range(15, 498, 11) for x in range(15, 498, 11)]
    q = poi_queue(pts)
    q["deformation"][:, 0] = 3
    q["deformation"][:, 6] = -2
    q["result"][::7, 2] = -4.0
    q["deformation"][3::11, 0] = np.nan
    q["result"][5::13, 2] = np.nan
    g = run_icgn_gpu(ref, tar, 15, q, sampling=sampling)
    o = q.copy()
    oracle.icgn(ref.astype(np.float32), tar.astype(np.float32), 15, o)
    cmp = compare(o, g)
    report(f"icgn edge cases ({sampling})", cmp)
    assert cmp["status_mismatch"] == 0
    assert cmp["pass"], cmp
    zc = o["result"][:, 2]
    assert (zc == -5).any() and (zc == -3).any() and (zc == -4).any()
    bad = ~finite_valid(o)
    # rejected POIs: every field but zncc equals the oracle's
    for name in ("x", "y", "deformation", "strain"):
        np.testing.assert_array_equal(g[name][bad], o[name][bad])


def test_icgn_self_adaptive_mixed_radii(oracle, c0):
    ref, tar = c0.ref, c0.tar
    q = fa.grid_queue(512, 512, 20, 12)
    rr = np.array([10, 15, 20], dtype=np.float32)[np.arange(len(q)) % 3]
    q["subset_radius"][:, 0] = q["subset_radius"][:, 1] = rr
    oracle.fftcc(ref.astype(np.float32), tar.astype(np.float32), 15, q)
    g = run_icgn_gpu(ref, tar, 15, q, self_adaptive=True)
    o = q.copy()
    oracle.icgn(ref.astype(np.float32), tar.astype(np.float32), 15, o, self_adaptive=True)
    cmp = compare(o, g)
    report("icgn mixed radii", cmp)
    assert cmp["pass"], cmp
    np.testing.assert_array_equal(g["subset_radius"], o["subset_radius"])


@pytest.mark.parametrize("sampling", ["auto", "pair"])
def test_icgn_iteration_cap_and_conv(oracle, c0, sampling):
    """A tight criterion with a small cap produces -4 the same way."""
    ref, tar = c0.ref, c0.tar
    q = c0.pois[:600].copy()
    oracle.fftcc(ref.astype(np.float32), tar.astype(np.float32), 15, q)
    g = run_icgn_gpu(ref, tar, 15, q, conv=1e-6, stop=2, sampling=sampling)
    o = q.copy()
    oracle.icgn(ref.astype(np.float32), tar.astype(np.float32), 15, o, conv=1e-6, stop=2)
    cmp = compare(o, g)
    report("icgn cap", cmp)
    assert (o["result"][:, 2] == -4).sum() > 100
    assert cmp["status_mismatch"] == 0


def test_icgn_table_sampling_parity(oracle, c0):
    """Every sample from the B-spline table (the fallback path) against the
    oracle, and against the default LDS-tile path, which evaluates the same
    polynomial from raw pixels."""
    ref, tar = c0.ref, c0.tar
    seeds = c0.pois.copy()
    oracle.fftcc(ref.astype(np.float32), tar.astype(np.float32), 15, seeds)
    g_tab = run_icgn_gpu(ref, tar, 15, seeds, sampling="table")
    g_lds = run_icgn_gpu(ref, tar, 15, seeds)
    o = seeds.copy()
    oracle.icgn(ref.astype(np.float32), tar.astype(np.float32), 15, o)
    cmp = compare(o, g_tab)
    report("icgn table sampling vs oracle", cmp)
    assert cmp["pass"], cmp
    both = compare(g_tab, g_lds)
    report("icgn table vs LDS tile", both)
    assert both["status_mismatch"] == 0 and max(both["max_du"], both["max_dv"]) < 1e-3


@pytest.mark.parametrize("sampling", ["auto", "pair"])
def test_icgn_large_strain_uses_table_fallback(oracle, sampling):
    """20 % strain: the warped subset outgrows the LDS tile, so iterations fall
    back to the table (pair sampling: to the target in device memory); parity
    holds."""
    sc = synth.affine_scene("strain", 384, 384, 3400, 7, t=(1.2, -0.7), grad=(0.2, 0.0, 0.0, 0.2),
                            radius=15, grid=16)
    q = sc.pois.copy()
    u, v, ux, uy, vx, vy = sc.truth(q["x"], q["y"])
    q["deformation"][:, 0] = np.round(u)
    q["deformation"][:, 6] = np.round(v)
    q["deformation"][:, 1] = 0.2
    q["deformation"][:, 8] = 0.2
    g = run_icgn_gpu(sc.ref, sc.tar, 15, q, sampling=sampling)
    o = q.copy()
    oracle.icgn(sc.ref.astype(np.float32), sc.tar.astype(np.float32), 15, o)
    cmp = compare(o, g)
    report("icgn 20% strain", cmp)
    assert cmp["pass"], cmp
    assert finite_valid(g).mean() > 0.5


@pytest.mark.parametrize("sampling", ["auto", "pair"])
def test_icgn_far_seed_restages_tile(oracle, c0, sampling):
    """Seeds 2.4 px from the solution move the subset past the tile's margin
    during the iterations, so the tile is staged again (pair sampling: the
    iteration reads the target in device memory); parity holds."""
    ref, tar = c0.ref, c0.tar
    q = c0.pois.copy()
    u, v, *_ = c0.truth(q["x"], q["y"])
    q["deformation"][:, 0] = u + 2.4
    q["deformation"][:, 6] = v - 2.4
    g = run_icgn_gpu(ref, tar, 15, q, stop=20, sampling=sampling)
    o = q.copy()
    oracle.icgn(ref.astype(np.float32), tar.astype(np.float32), 15, o, stop=20)
    cmp = compare(o, g)
    report("icgn far seeds", cmp)
    assert cmp["pass"], cmp
    moved = finite_valid(g) & (np.abs(g["deformation"][:, 0] - q["deformation"][:, 0]) > 2.0)
    assert moved.sum() > 100


def test_icgn_deterministic(c0):
    seeds = c0.pois.copy()
    seeds["deformation"][:, 0] = 2.0
    seeds["deformation"][:, 6] = -2.0
    a = run_icgn_gpu(c0.ref, c0.tar, 15, seeds)
    b = run_icgn_gpu(c0.ref, c0.tar, 15, seeds)
    assert fields_equal(a, b)


# ------------------------------------------------------------ frame driver --

def test_frame_driver_matches_engines(oracle, c0):
    say("frame driver config0 first frame")
    with fa.FrameDriver(15) as d:
        d.set_reference(fa.Image2D(c0.ref))
        d.set_pois(c0.pois)
        res, st = d.process(fa.Image2D(c0.tar), policy=0)
    q = c0.pois.copy()
    with fa.FFTCC2D(15, 15) as f:
        f.setImages(fa.Image2D(c0.ref), fa.Image2D(c0.tar))
        f.prepareRef(q)
        f.compute(q)
    g = run_icgn_gpu(c0.ref, c0.tar, 15, q)
    assert fields_equal(res, g)
    assert st["fftcc_pois"] == len(q)
    assert st["valid_count"] == int(finite_valid(res).sum())
    say("frame stats", st)


def test_frame_driver_sequence(oracle, c0):
    """Frames of growing translation against a fixed reference, with the
    reference's default policy (skip-FFTCC seeding, bad-ratio guard, POI-level
    repair); each frame is checked against the CPU restatement of the policy."""
    from tests.policy import PolicyOracle
    ref = c0.ref
    steps = [(0.0, 0.0), (0.6, -0.4), (1.3, -0.9), (6.0, 4.0), (6.4, 4.3)]
    po = PolicyOracle(oracle, ref, c0.pois, 15)
    with fa.FrameDriver(15) as d:
        d.set_reference(fa.Image2D(ref))
        d.set_pois(c0.pois)
        for k, t in enumerate(steps):
            sc = synth.affine_scene("seq", 512, 512, 6000, 20260, t=t,
                                    grad=(4e-4 * k, 0, 0, 4e-4 * k), radius=15, grid=10)
            res, st = d.process(fa.Image2D(sc.tar))
            o, ost = po.frame(sc.tar)
            say("frame", k, {x: st[x] for x in ("skip_fftcc_used", "fallback_used", "poi_fallback_count",
                                                  "bad_ratio", "valid_ratio", "iteration_mean")}, ost)
            for key in ("skip_fftcc_used", "fallback_used", "poi_fallback_count"):
                assert st[key] == ost[key], key
            cmp = compare(o, res)
            report(f"frame {k}", cmp)
            assert cmp["pass"], cmp


# ------------------------------------------------------------- full size --

def test_config1_full_size_properties(oracle):
    """2048^2, ~162k POIs through the frame driver; parity on a POI sample
    against the oracle, accuracy against the known field on every POI."""
    sc = synth.config1()
    say("config1", sc.ref.shape, len(sc.pois), "POIs", sc.meta["initial_guess"])
    with fa.FrameDriver(15) as d:
        d.set_reference(fa.Image2D(sc.ref))
        d.set_pois(sc.pois)
        res, st = d.process(fa.Image2D(sc.tar), policy=0)
    say("config1 stats", st)
    ok = finite_valid(res)
    assert ok.mean() > 0.99
    u, v, *_ = sc.truth(res["x"], res["y"])
    err = np.maximum(np.abs(res["deformation"][ok, 0] - u[ok]), np.abs(res["deformation"][ok, 6] - v[ok]))
    say("config1 truth p99", float(np.percentile(err, 99)), "max", float(err.max()))
    assert np.percentile(err, 99) < 0.1
    rng = np.random.default_rng(0)
    pick = np.sort(rng.choice(len(sc.pois), 3000, replace=False))
    ref32, tar32 = sc.ref.astype(np.float32), sc.tar.astype(np.float32)
    # FFTCC at full size: bit-identical on the sample
    o = sc.pois[pick].copy()
    oracle.fftcc(ref32, tar32, 15, o)
    g = sc.pois[pick].copy()
    with fa.FFTCC2D(15, 15) as f:
        f.setImages(fa.Image2D(sc.ref), fa.Image2D(sc.tar))
        f.compute(g)
    assert fields_equal(g, o)
    oracle.icgn(ref32, tar32, 15, o)
    cmp = compare(o, res[pick], max_iteration_mismatch=0.01)
    report("config1 sample", cmp)
    assert cmp["pass"], cmp
    o64 = sc.pois[pick].copy()
    oracle.fftcc(ref32, tar32, 15, o64)
    oracle.icgn(ref32, tar32, 15, o64, f64=True)
    c64 = compare(o64, res[pick], max_iteration_mismatch=0.01)
    report("config1 sample vs FP64 oracle", c64)
    assert c64["pass"], c64


# ------------------------------------------------------------ learned seed --

def _learned_frame(sc, params, threshold=0.8):
    from fasticgn_amd import seeder as sd
    with sd.LearnedSeeder(params) as s, fa.FrameDriver(sc.radius) as d:
        d.set_reference(fa.Image2D(sc.ref))
        d.set_pois(sc.pois)
        d.set_seeder(s, threshold)
        d.upload_target(fa.Image2D(sc.tar))
        st = d.run(fa.driver.LEARNED_SEED)
        return d.download(), st


def test_learned_seed_frame_parity(oracle):
    """The shipped seeder on the 2048^2 frame (POIs start from the global
    integer guess): same parity gate against the CPU oracle as the FFTCC-seeded
    path on a POI sample, validity flags included, and accuracy against the
    known field on every POI."""
    from fasticgn_amd import seeder as sd
    assert sd.WEIGHTS_PATH.exists()
    sc = synth.config1()
    res, st = _learned_frame(sc, None)
    say("learned config1 stats", {k: st[k] for k in ("seed_pois", "seed_fallback_count", "valid_count",
                                                     "iteration_mean", "seed_ms", "icgn_ms")})
    assert st["seed_pois"] == len(sc.pois)
    assert st["seed_fallback_count"] < 0.05 * len(sc.pois)
    ok = finite_valid(res)
    assert ok.mean() > 0.99
    u, v, *_ = sc.truth(res["x"], res["y"])
    err = np.maximum(np.abs(res["deformation"][ok, 0] - u[ok]), np.abs(res["deformation"][ok, 6] - v[ok]))
    assert np.percentile(err, 99) < 0.1
    rng = np.random.default_rng(1)
    pick = np.sort(rng.choice(len(sc.pois), 3000, replace=False))
    o = sc.pois[pick].copy()
    ref32, tar32 = sc.ref.astype(np.float32), sc.tar.astype(np.float32)
    oracle.fftcc(ref32, tar32, 15, o)
    oracle.icgn(ref32, tar32, 15, o)
    cmp = compare(o, res[pick])
    report("learned config1 sample", cmp)
    assert cmp["pass"], cmp


def test_learned_seed_bad_seeds_fall_back_to_fftcc(oracle, c0):
    """A seeder whose u output is off by 40 px: every IC-GN from its seeds
    fails or lands on a low ZNCC, FFTCC redoes those POIs, and the frame equals
    the plain FFTCC-seeded frame except where the redo is out of bounds (ZNCC
    -3, which never replaces: shouldReplacePoiWithFallback).  The POIs the
    consensus leaves unseeded take the redo as it is."""
    from fasticgn_amd import seeder as sd
    p = np.load(sd.WEIGHTS_PATH).copy()
    p[-6] += 40.0  # b5[0]: u seed + 40 px
    res, st = _learned_frame(c0, p)
    with fa.FrameDriver(15) as d:
        d.set_reference(fa.Image2D(c0.ref))
        d.set_pois(c0.pois)
        plain, _ = d.process(fa.Image2D(c0.tar), policy=0)
    say("bad seeder stats", {k: st[k] for k in ("seed_pois", "seed_fallback_count", "valid_count")})
    assert st["seed_fallback_count"] >= 0.95 * len(c0.pois)
    same = np.all(res["deformation"] == plain["deformation"], axis=1) & np.all(res["result"] == plain["result"], axis=1)
    assert same.mean() >= 0.95
    assert np.all(plain["result"][~same, 2] == -3) and np.all(res["result"][~same, 2] == -3)
    cmp = compare(plain, res)
    assert cmp["pass"], cmp


def test_learned_seed_two_passes_parity(oracle, c0):
    """Config 0 starts every POI at guess 0 with u = 2.35, v = -1.70: a second
    seeder pass (crops re-centred on the first pass's rounded u, v) keeps the
    parity gate against the CPU oracle and needs no more IC-GN iterations."""
    from fasticgn_amd import seeder as sd
    out = {}
    for passes in (1, 2):
        with sd.LearnedSeeder() as s, fa.FrameDriver(15) as d:
            d.set_reference(fa.Image2D(c0.ref))
            d.set_pois(c0.pois)
            d.set_seeder(s, 0.8)
            d.set_seed_passes(passes)
            out[passes] = d.process(fa.Image2D(c0.tar), policy=fa.driver.LEARNED_SEED)
    o = c0.pois.copy()
    ref32, tar32 = c0.ref.astype(np.float32), c0.tar.astype(np.float32)
    oracle.fftcc(ref32, tar32, 15, o)
    oracle.icgn(ref32, tar32, 15, o)
    cmp = compare(o, out[2][0])
    report("learned 2 passes config0", cmp)
    assert cmp["pass"], cmp
    say("iterations 1 pass", out[1][1]["iteration_sum"], "2 passes", out[2][1]["iteration_sum"])
    assert out[2][1]["iteration_sum"] <= out[1][1]["iteration_sum"]
    with fa.FrameDriver(15) as d:
        with pytest.raises(ValueError):
            d.set_seed_passes(0)
