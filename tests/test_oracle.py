"""Pins the CPU oracle (oracle/oracle.c) to the reference's own known answers,
to FP64 brute force and to closed forms.  CPU only.

Each check is chosen so a plausible restatement mistake fails it: a wrong tap
sign or order in the gradient, a transposed BC product, a conj on the wrong
operand or a dropped Nyquist term in the DFT, a wrong wrap of the peak, a
wrong pivot in the 6x6 inverse, a wrong sign in the inverse-compositional
update.
"""
import math

import numpy as np
import pytest

from fasticgn_amd import synth
from fasticgn_amd.opencorr import poi_queue


def brute_circular(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """c[v][u] = n^2 * sum_{y,x} a[y][x] * b[(y+v)%n][(x+u)%n], in FP64 by
    direct summation over all lags (what FFTW's unnormalised c2r of
    conj(R)*T equals)."""
    n = a.shape[0]
    A = a.astype(np.float64)
    B = b.astype(np.float64)
    out = np.empty((n, n))
    for v in range(n):
        for u in range(n):
            out[v, u] = np.sum(A * np.roll(B, (-v, -u), axis=(0, 1)))
    return out * n * n


# ---------------------------------------------------------------- FFTCC ----

def test_lcg_known_answer(oracle):
    """smoke/library_boundary_probe.cpp:41-62: every POI u = 2, v = -1, zncc > 0.5."""
    ref, tar, q = synth.lcg_pair()
    oracle.fftcc(ref, tar, 7, q)
    assert np.all(q["deformation"][:, 0] == 2.0)
    assert np.all(q["deformation"][:, 6] == -1.0)
    assert np.all(q["result"][:, 2] > 0.5)
    assert np.all(q["result"][:, 0] == 0.0) and np.all(q["result"][:, 1] == 0.0)


def test_lcg_known_answer_cached(oracle):
    ref, tar, q = synth.lcg_pair()
    prep = oracle.PreparedRef(ref, 7, q)
    prep.compute(tar, q)
    assert np.all(q["deformation"][:, 0] == 2.0) and np.all(q["deformation"][:, 6] == -1.0)


@pytest.mark.parametrize("n", [2, 4, 6, 8, 10, 12, 14, 18, 20, 22, 26, 28, 30, 36, 40, 42, 50, 56, 60, 62, 64])
def test_dft_correlation_matches_brute_force(oracle, n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n)).astype(np.float32)
    b = (np.roll(a, (min(3, n - 1), -min(5, n - 1)), axis=(0, 1))
         + 0.2 * rng.normal(size=(n, n))).astype(np.float32)
    a -= a.mean(dtype=np.float64).astype(np.float32)
    b -= b.mean(dtype=np.float64).astype(np.float32)
    got = oracle.correlate(a, b)
    want = brute_circular(a, b)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 2e-5 * scale
    assert np.argmax(got) == np.argmax(want)
    # swapping the operands mirrors the lag: the peak moves unless it is at 0
    swapped = oracle.correlate(b, a)
    assert np.abs(swapped - want).max() > 1e-3 * scale or n <= 2


def test_twiddles_closed_form(oracle):
    for n in (2, 14, 30, 40, 64):
        tw = oracle.twiddles(n).reshape(n, 2)
        k = np.arange(n)
        np.testing.assert_array_equal(tw[:, 0], np.cos(2 * np.pi * k / n).astype(np.float32))
        np.testing.assert_array_equal(tw[:, 1], (-np.sin(2 * np.pi * k / n)).astype(np.float32))


def test_identical_windows_give_unit_zncc(oracle):
    sc = synth.config0()
    ref = sc.ref.astype(np.float32)
    q = poi_queue([(100, 100), (200, 300), (400, 50)])
    oracle.fftcc(ref, ref, 15, q)
    assert np.all(q["deformation"][:, 0] == 0) and np.all(q["deformation"][:, 6] == 0)
    np.testing.assert_allclose(q["result"][:, 2], 1.0, atol=2e-6)


def periodic_pair(r: int, dx: int, dy: int, seed: int = 7):
    """Random image with period 2r in both axes and its copy moved by
    (dx, dy): every 2r x 2r window of the target is an exact circular shift of
    the reference window, so the correlation peak sits exactly at the lag."""
    n = 2 * r
    rng = np.random.default_rng(seed)
    tile = rng.integers(0, 256, size=(n, n)).astype(np.float32)
    ref = np.tile(tile, (6, 6))
    tar = np.roll(ref, (dy, dx), axis=(0, 1))
    return ref, tar


@pytest.mark.parametrize("r", [7, 15])
def test_wrap_rule_full_capture_range(oracle, r):
    """Lags above r wrap to negative (src/oc_fftcc.cpp:376-386): after the
    wrap the result covers exactly [-(r - 1), r] on both axes."""
    pts = [(3 * r, 3 * r), (5 * r, 4 * r)]
    for d in list(range(-(r - 1), r + 1)):
        for dx, dy in ((d, 0), (0, d), (d, -d if -d > -r else 0)):
            ref, tar = periodic_pair(r, dx, dy)
            q = poi_queue(pts)
            oracle.fftcc(ref, tar, r, q)
            assert np.all(q["deformation"][:, 0] == dx) and np.all(q["deformation"][:, 6] == dy), (dx, dy)
            np.testing.assert_allclose(q["result"][:, 2], 1.0, atol=1e-5)


@pytest.mark.parametrize("shift", [(-3, 0), (3, 0), (0, -3), (0, 3), (2, -3), (-3, 2)])
def test_small_shift_on_speckle(oracle, shift):
    r = 15
    sc = synth.config0()
    ref = sc.ref.astype(np.float32)
    dx, dy = shift
    tar = np.roll(ref, (dy, dx), axis=(0, 1))
    q = poi_queue([(256, 256), (200, 220), (300, 100), (100, 300)])
    oracle.fftcc(ref, tar, r, q)
    assert np.all(q["deformation"][:, 0] == dx) and np.all(q["deformation"][:, 6] == dy)


def test_boundary_pois_untouched(oracle):
    """compute(POI2D*) returns without touching a POI whose reference or
    target window leaves [r, size - r) (src/oc_fftcc.cpp:404-412)."""
    sc = synth.config0()
    ref = sc.ref.astype(np.float32)
    q = poi_queue([(14, 100), (100, 497), (15, 15), (496, 496)])
    q["result"][:, 2] = -7.0
    q["deformation"][2, 0] = -1.0  # target window at x = 14 -> outside
    before = q.copy()
    oracle.fftcc(ref, ref, 15, q)
    assert q[0].tobytes() == before[0].tobytes()
    assert q[1].tobytes() == before[1].tobytes()
    assert q[2].tobytes() == before[2].tobytes()
    assert q[3]["result"][2] != -7.0  # (496, 496): 496 < 512 - 15, inside


def test_cached_equals_uncached_on_integer_pois(oracle):
    sc = synth.config0()
    ref, tar = sc.ref.astype(np.float32), sc.tar.astype(np.float32)
    q1 = sc.pois.copy()
    q2 = sc.pois.copy()
    oracle.fftcc(ref, tar, 15, q1)
    oracle.PreparedRef(ref, 15, q2).compute(tar, q2)
    assert q1.tobytes() == q2.tobytes()


# ------------------------------------------------------------- gradient ----

def test_gradient_taps_and_border(oracle):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, size=(40, 50)).astype(np.float32)
    gx, gy = oracle.gradient(img)
    f = img.astype(np.float64)
    want_x = np.zeros_like(f)
    want_x[:, 2:-2] = (-f[:, 4:] + 8 * f[:, 3:-1] - 8 * f[:, 1:-3] + f[:, :-4]) / 12
    want_y = np.zeros_like(f)
    want_y[2:-2, :] = (-f[4:, :] + 8 * f[3:-1, :] - 8 * f[1:-3, :] + f[:-4, :]) / 12
    np.testing.assert_allclose(gx, want_x, atol=1e-4)
    np.testing.assert_allclose(gy, want_y, atol=1e-4)
    assert np.all(gx[:, :2] == 0) and np.all(gx[:, -2:] == 0)
    assert np.all(gy[:2, :] == 0) and np.all(gy[-2:, :] == 0)


def test_gradient_exact_on_cubics(oracle):
    """The 4-tap stencil differentiates polynomials of degree <= 4 exactly."""
    y, x = np.mgrid[0:24, 0:24].astype(np.float64)
    f = 0.01 * x ** 3 - 0.3 * x * x + 2 * x + 0.02 * y ** 3 + y
    gx, gy = oracle.gradient(f.astype(np.float32), f64=True)
    dfx = 0.03 * x * x - 0.6 * x + 2
    dfy = 0.06 * y * y + 1
    np.testing.assert_allclose(gx[:, 2:-2], dfx[:, 2:-2], atol=1e-4)
    np.testing.assert_allclose(gy[2:-2, :], dfy[2:-2, :], atol=1e-4)


# ------------------------------------------------------------- B-spline ----

BC = np.array([[-144, 384, -384, 144], [342, -702, 450, -90], [-198, -18, 270, -54],
               [0, 336, 0, 0]], dtype=np.float64) / 336.0


def test_bspline_table_is_bc_q_bct(oracle):
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, size=(12, 14)).astype(np.float32)
    coef = oracle.bspline(img, f64=True)
    for r, c in [(1, 1), (5, 7), (8, 10), (9, 11)]:
        Q = img[r - 1:r + 3, c - 1:c + 3].astype(np.float64)
        P = BC @ Q @ BC.T
        want = P[::-1, ::-1].ravel()
        np.testing.assert_allclose(coef[r, c], want, rtol=1e-12, atol=1e-9)
    assert np.all(coef[0] == 0) and np.all(coef[:, 0] == 0)
    assert np.all(coef[-2:] == 0) and np.all(coef[:, -2:] == 0)


def test_bspline_constant_and_nodes(oracle):
    img = np.full((10, 10), 73.0, dtype=np.float32)
    coef = oracle.bspline(img)
    np.testing.assert_allclose(coef[3, 4, 0], 73.0, rtol=1e-6)
    assert np.abs(coef[3, 4, 1:]).max() < 1e-4
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, size=(16, 16)).astype(np.float32)
    coef = oracle.bspline(img)
    for (x, y) in [(1, 1), (5, 9), (12, 13)]:
        assert oracle.bspline_eval(coef, float(x), float(y)) == img[y, x]
    assert oracle.bspline_eval(coef, 0.999, 5.0) == -1.0
    assert oracle.bspline_eval(coef, 5.0, 14.0) == -1.0  # y >= h - 2


def test_bspline_eval_polynomial(oracle):
    rng = np.random.default_rng(11)
    img = rng.normal(size=(10, 10)).astype(np.float32)
    coef = oracle.bspline(img, f64=True)
    x, y = 4.3, 5.6
    xi, yi = 4, 5
    dx, dy = x - xi, y - yi
    P = BC @ img[yi - 1:yi + 3, xi - 1:xi + 3].astype(np.float64) @ BC.T
    want = np.array([dy ** 3, dy ** 2, dy, 1.0]) @ P @ np.array([dx ** 3, dx ** 2, dx, 1.0])
    assert abs(oracle.bspline_eval(coef, x, y) - want) < 1e-10


# ------------------------------------------------------------ inverse6 ----

def test_inverse6_spd_and_pivoting(oracle):
    rng = np.random.default_rng(2)
    for _ in range(5):
        M = rng.normal(size=(6, 6))
        H = M @ M.T + 0.5 * np.eye(6)
        np.testing.assert_allclose(oracle.inverse6(H, f64=True), np.linalg.inv(H), rtol=1e-9, atol=1e-12)
        inv32 = oracle.inverse6(H.astype(np.float32)).astype(np.float64)
        np.testing.assert_allclose(inv32 @ H, np.eye(6), atol=2e-3)
    P = np.eye(6)[[3, 0, 5, 1, 4, 2]] * np.array([1, 2, 3, 4, 5, 6.0])[:, None]
    np.testing.assert_allclose(oracle.inverse6(P, f64=True), np.linalg.inv(P), atol=1e-14)


def test_inverse6_singular_gives_nonfinite(oracle):
    H = np.zeros((6, 6))
    H[:3, :3] = np.eye(3)
    assert not np.all(np.isfinite(oracle.inverse6(H, f64=True)))


# ---------------------------------------------------------------- ICGN ----

def shifted(ref: np.ndarray, dx: int, dy: int, fill: float = 40.0) -> np.ndarray:
    tar = np.full_like(ref, fill)
    h, w = ref.shape
    tar[max(dy, 0):h + min(dy, 0), max(dx, 0):w + min(dx, 0)] = \
        ref[max(-dy, 0):h - max(dy, 0), max(-dx, 0):w - max(dx, 0)]
    return tar


def test_icgn_integer_shift_one_step(oracle):
    """smoke/run_cache_regression.py:22-28: an integer shift of noise-free
    speckle is recovered to 0.01 px from the FFTCC seed."""
    sc = synth.config0()
    ref = sc.ref.astype(np.float32)
    tar = shifted(ref, 5, -3)
    q = poi_queue([(x, y) for y in range(60, 450, 40) for x in range(60, 450, 40)])
    oracle.fftcc(ref, tar, 15, q)
    oracle.icgn(ref, tar, 15, q)
    assert np.all(q["result"][:, 2] > 0.99)
    assert np.abs(q["deformation"][:, 0] - 5).max() < 0.01
    assert np.abs(q["deformation"][:, 6] + 3).max() < 0.01
    assert np.median(q["result"][:, 3]) == 1  # exact seeds stop after one step


def test_icgn_affine_truth_and_fp64_twin(oracle):
    sc = synth.config0()
    ref, tar = sc.ref.astype(np.float32), sc.tar.astype(np.float32)
    q = sc.pois.copy()
    oracle.fftcc(ref, tar, 15, q)
    q64 = q.copy()
    oracle.icgn(ref, tar, 15, q)
    oracle.icgn(ref, tar, 15, q64, f64=True)
    ok = q["result"][:, 2] >= 0
    assert ok.mean() > 0.97
    u, v, ux, uy, vx, vy = sc.truth(q["x"], q["y"])
    eu = q["deformation"][ok, 0] - u[ok]
    ev = q["deformation"][ok, 6] - v[ok]
    # accuracy scale of smoke/run_subpixel_accuracy_gate.py LIMITS (s15 rmse 0.030)
    assert np.sqrt(np.mean(eu ** 2)) < 0.03 and np.sqrt(np.mean(ev ** 2)) < 0.03
    np.testing.assert_array_equal(q["result"][:, 2] >= 0, q64["result"][:, 2] >= 0)
    assert np.abs(q["deformation"][ok, 0] - q64["deformation"][ok, 0]).max() < 1e-3
    assert np.abs(q["deformation"][ok, 1] - ux[ok]).mean() < 2e-3


def test_icgn_oob_band_closed_form(oracle):
    """smoke/run_oob_rejection_gate.py:53-56: with a whole-pixel shift the
    rejected set is x + dx - r < 1 or y + dy - r < 1 or x + dx + r >= w - 2 or
    y + dy + r >= h - 2."""
    sc = synth.config0()
    ref = sc.ref.astype(np.float32)
    dx, dy, r = 3, -2, 15
    tar = shifted(ref, dx, dy)
    h, w = ref.shape
    q = poi_queue([(x, y) for y in range(15, h - 15, 7) for x in range(15, w - 15, 7)])
    q["deformation"][:, 0] = dx
    q["deformation"][:, 6] = dy
    oracle.icgn(ref, tar, r, q)
    x, y = q["x"], q["y"]
    pred = (x + dx - r < 1) | (y + dy - r < 1) | (x + dx + r >= w - 2) | (y + dy + r >= h - 2)
    got = q["result"][:, 2] == -3
    np.testing.assert_array_equal(got, pred)
    # the rejected POIs keep the incoming guess
    assert np.all(q["deformation"][got, 0] == dx)


def test_icgn_zero_texture_gives_minus5(oracle):
    """smoke/run_zero_texture_gate.py:9-35: a constant subset fails with a
    0/0 cost on the CPU path (-5), u, v keep the incoming guess."""
    sc = synth.config0()
    ref = sc.ref.astype(np.float32).copy()
    ref[100:200, 100:200] = 128.0
    q = poi_queue([(150, 150), (300, 300)])
    oracle.fftcc(ref, ref, 15, q)
    oracle.icgn(ref, ref, 15, q)
    assert q["result"][0, 2] == -5.0
    assert q["deformation"][0, 0] == 0.0 and q["deformation"][0, 6] == 0.0
    assert q["result"][1, 2] > 0.999


def test_icgn_entry_gate(oracle):
    sc = synth.config0()
    ref = sc.ref.astype(np.float32)
    q = poi_queue([(100, 100), (100, 100), (100, 100), (10, 100), (100, 100)])
    q["result"][0, 2] = -4.0           # negative incoming zncc is kept
    q["deformation"][1, 0] = np.nan    # NaN guess -> -3
    q["deformation"][2, 6] = 600.0     # |v| >= height -> -3
    q["result"][4, 2] = np.nan         # NaN zncc passes the gate (src/oc_icgn.cpp:187-190)
    oracle.icgn(ref, ref, 15, q)
    assert q["result"][0, 2] == -4.0
    assert q["result"][1, 2] == -3.0
    assert q["result"][2, 2] == -3.0
    assert q["result"][3, 2] == -3.0   # subset leaves the image
    assert q["result"][4, 2] > 0.999


def test_icgn_self_adaptive_radius(oracle):
    sc = synth.config0()
    ref, tar = sc.ref.astype(np.float32), sc.tar.astype(np.float32)
    q = sc.pois[200:260].copy()
    q["subset_radius"][:, 0] = q["subset_radius"][:, 1] = np.array([10, 15, 20] * 20, dtype=np.float32)
    q["deformation"][:, 0] = 2.0
    q["deformation"][:, 6] = -2.0
    oracle.icgn(ref, tar, 15, q, self_adaptive=True)
    ok = q["result"][:, 2] >= 0
    assert ok.all()
    np.testing.assert_array_equal(q["subset_radius"][:, 0], [10, 15, 20] * 20)


def _offset_scene():
    return synth.affine_scene("offset", 384, 384, 3400, 11, t=(1.3, -0.8), grad=(0.04, 0.02, -0.03, 0.05),
                              radius=15, grid=24)


def test_icgn_center_offset_closed_form(oracle):
    """compute(queue, center_offset_queue) (src/oc_icgn.cpp:399-625): the warp
    acts about the POI plus its offset, so on an affine field the converged u, v
    are the displacement at (x + ox, y + oy) (a wrong sign lands 2 |grad . o|
    away) and the gradients are the field's; zero offsets reproduce
    compute(queue) bit for bit, in FP32 and in the FP64 twin."""
    sc = _offset_scene()
    ref, tar = sc.ref.astype(np.float32), sc.tar.astype(np.float32)
    q = sc.pois.copy()
    u, v, ux, uy, vx, vy = sc.truth(q["x"], q["y"])
    d = q["deformation"]
    d[:, 0], d[:, 6] = np.round(u), np.round(v)
    d[:, 1], d[:, 2], d[:, 7], d[:, 8] = ux, uy, vx, vy
    for f64 in (False, True):
        plain, zero = q.copy(), q.copy()
        oracle.icgn(ref, tar, 15, plain, f64=f64)
        oracle.icgn(ref, tar, 15, zero, f64=f64, center_offsets=np.zeros((len(q), 2), np.float32))
        assert plain.tobytes() == zero.tobytes()
    rng = np.random.default_rng(5)
    off = rng.uniform(-3, 3, (len(q), 2)).astype(np.float32)
    g = q.copy()
    oracle.icgn(ref, tar, 15, g, center_offsets=off)
    ok = g["result"][:, 2] >= 0
    assert ok.mean() > 0.8
    ut, vt, *_ = sc.truth(g["x"] + off[:, 0], g["y"] + off[:, 1])
    uw, vw, *_ = sc.truth(g["x"] - off[:, 0], g["y"] - off[:, 1])
    err = np.hypot(g["deformation"][ok, 0] - ut[ok], g["deformation"][ok, 6] - vt[ok])
    wrong = np.hypot(g["deformation"][ok, 0] - uw[ok], g["deformation"][ok, 6] - vw[ok])
    assert np.percentile(err, 99) < 0.03
    assert np.median(wrong) > 0.1
    assert np.abs(g["deformation"][ok, 1] - ux[ok]).mean() < 2e-3
    assert np.abs(g["deformation"][ok, 8] - vy[ok]).mean() < 2e-3
    # the plain run measures the displacement at the POI itself
    p = q.copy()
    oracle.icgn(ref, tar, 15, p)
    okp = ok & (p["result"][:, 2] >= 0)
    assert np.abs(p["deformation"][okp, 0] - u[okp]).max() < 0.05


def test_fftcc_window_sum_order_reference_vs_device(oracle):
    """FFTCC's window mean and norm in the reference's order (one serial sum
    in row order, src/oc_fftcc.cpp:280-299, 415-437) and in the device's
    (row partials, which the kernel reproduces bit for bit): the integer
    peaks are identical on the 512^2 u8 scene and a 12-bit scene, and the
    ZNCC differs by FP32 summation noise only (measured 8.7e-6 and 1.8e-6
    relative)."""
    from fasticgn_amd import synth
    c0 = synth.config0()
    c2 = synth.config2(size=256, frames=500)
    for ref, tar, depth in ((c0.ref, c0.tar, "native"), (c2.ref, c2.frame(120), "right")):
        pois = c0.pois if depth == "native" else c2.pois
        ref32, tar32 = oracle.as_f32(ref, depth), oracle.as_f32(tar, depth)
        dev, ser = pois.copy(), pois.copy()
        oracle.fftcc(ref32, tar32, 15, dev)
        oracle.fftcc(ref32, tar32, 15, ser, serial_sums=True)
        assert np.array_equal(dev["deformation"], ser["deformation"])
        zd, zs = dev["result"][:, 2], ser["result"][:, 2]
        assert (zd != zs).any()  # the orders do differ in the last bits
        assert np.max(np.abs(zd - zs) / np.abs(zs)) < 2e-5
