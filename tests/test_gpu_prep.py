"""The preparation stages on the device, checked directly against the oracle
(through fi_icgn_export), not only through the IC-GN results they feed.

  Gradient2D4        src/oc_gradient.cpp:37-79      bit-identical to or_gradient_f32
  BicubicBspline     src/oc_cubic_bspline.cpp:84-132 table bit-identical to
                                                     or_bspline_prepare_f32
  per-POI records    src/oc_icgn.cpp:198-234         mean, norm and sum (ref - mean)
                                                     against FP64 sums; the inverse
                                                     Hessian against an FP64 Hessian
                                                     and inverse within a bound set
                                                     by the Hessian's condition number
  12-bit input       src/oc_image.cpp:163-185        left-aligned u16 (value x 16)
                                                     gives the same results as the
                                                     right-aligned image, bit for bit
"""
import numpy as np
import pytest

import fasticgn_amd as fa
from fasticgn_amd import synth
from fasticgn_amd.opencorr import poi_queue

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c0():
    return synth.config0()


@pytest.fixture(scope="module")
def u16_scene():
    seq = synth.config2(size=256, frames=500)
    return seq.ref, seq.frame(120)


def engine(ref_img, tar_img, r=15):
    ic = fa.ICGN2D1(r, r, 0.001, 10)
    ic.setImages(ref_img, tar_img)
    ic.prepare()
    return ic


@pytest.mark.parametrize("kind", ["u8", "u16"])
def test_gradients_and_table_bit_identical(oracle, c0, u16_scene, kind):
    if kind == "u8":
        ref, tar, depth = c0.ref, c0.tar, "native"
    else:
        (ref, tar), depth = u16_scene, "right"
    h, w = ref.shape
    ref32, tar32 = oracle.as_f32(ref, depth), oracle.as_f32(tar, depth)
    with engine(fa.Image2D(ref, depth12=depth), fa.Image2D(tar, depth12=depth)) as ic:
        got = {k: ic.export(k, (h, w)) for k in ("ref", "gx", "gy", "tar", "bspline")}
    gx, gy = oracle.gradient(ref32)
    coef = oracle.bspline(tar32)
    assert got["ref"].tobytes() == ref32.tobytes()
    assert got["tar"].tobytes() == tar32.tobytes()
    assert got["gx"].tobytes() == gx.tobytes(), np.abs(got["gx"] - gx).max()
    assert got["gy"].tobytes() == gy.tobytes(), np.abs(got["gy"] - gy).max()
    assert got["bspline"].tobytes() == coef.tobytes(), np.abs(got["bspline"] - coef).max()
    # the zero border of the stencil and of the table
    assert not got["gx"][:, :2].any() and not got["gx"][:, -2:].any() and not got["gy"][:2].any()
    assert not got["bspline"][0].any() and not got["bspline"][-2:].any()


def fp64_records(ref32, gx64, gy64, q, r):
    """Mean, norm, sum (ref - mean) and the Hessian of every POI in FP64
    (src/oc_icgn.cpp:198-234: subset at (int)(x - r), gradients at (int)x + local)."""
    out = []
    k = np.arange(-r, r + 1, dtype=np.float64)
    yl, xl = np.meshgrid(k, k, indexing="ij")
    for x, y in zip(q["x"], q["y"]):
        ulx, uly = int(x - r), int(y - r)
        sub = ref32[uly:uly + 2 * r + 1, ulx:ulx + 2 * r + 1].astype(np.float64)
        px, py = int(x), int(y)
        gxs = gx64[py - r:py + r + 1, px - r:px + r + 1]
        gys = gy64[py - r:py + r + 1, px - r:px + r + 1]
        sd = np.stack([gxs, gxs * xl, gxs * yl, gys, gys * xl, gys * yl]).reshape(6, -1)
        mean = sub.mean()
        out.append((mean, sub, sd @ sd.T))
    return out


def test_prep_records_against_fp64(oracle, c0):
    r = 15
    ref = c0.ref
    tar = c0.tar
    q = c0.pois[::7].copy()
    # two POIs the reference gate rejects (subset too close to the border)
    q = np.concatenate([q, poi_queue([(r - 1, 100), (200, 512 - r)])])
    with engine(fa.Image2D(ref), fa.Image2D(tar), r) as ic:
        g = q.copy()
        ic.compute(g)
        rec = ic.export("prep", n_records=len(q))
    ref32 = oracle.as_f32(ref)
    gx64, gy64 = oracle.gradient(ref32, f64=True)
    assert (rec["status"][-2:] == 0).all() and (rec["status"][:-2] == 1).all()
    assert (rec["rx"] == r).all() and (rec["ry"][:-2] == r).all()
    worst = {"mean_ulp": 0.0, "norm_rel": 0.0, "ref_sum": 0.0, "hinv_over_bound": 0.0}
    for i, (mean64, sub, H) in enumerate(fp64_records(ref32, gx64, gy64, q[:-2], r)):
        m = rec["mean"][i]
        ulp = float(np.spacing(np.float32(abs(mean64))))
        worst["mean_ulp"] = max(worst["mean_ulp"], abs(float(m) - mean64) / ulp)
        # norm and sum of the subset centred on the device's own FP32 mean
        d = sub - np.float64(m)
        norm64 = np.sqrt((d * d).sum())
        worst["norm_rel"] = max(worst["norm_rel"], abs(float(rec["norm"][i]) - norm64) / norm64)
        worst["ref_sum"] = max(worst["ref_sum"], abs(float(rec["ref_sum"][i]) - d.sum()) / np.abs(d).sum())
        Hi64 = np.linalg.inv(H)
        hi = rec["hinv"][i].reshape(6, 6).astype(np.float64)
        bound = 1e-6 * np.linalg.cond(H) + 2.0 ** -23
        err = np.abs(hi - Hi64).max() / np.abs(Hi64).max()
        worst["hinv_over_bound"] = max(worst["hinv_over_bound"], err / bound)
    say("prep records vs FP64", len(q) - 2, "POIs:", worst)
    assert worst["mean_ulp"] <= 2.0
    assert worst["norm_rel"] <= 1e-6
    assert worst["ref_sum"] <= 1e-6
    assert worst["hinv_over_bound"] <= 1.0


def test_prep_records_self_adaptive_radii(oracle, c0):
    q = fa.grid_queue(512, 512, 20, 37)
    rr = np.array([7, 12, 20], dtype=np.float32)[np.arange(len(q)) % 3]
    q["subset_radius"][:, 0] = q["subset_radius"][:, 1] = rr
    with engine(fa.Image2D(c0.ref), fa.Image2D(c0.tar)) as ic:
        ic.setSelfAdaptive(True)
        g = q.copy()
        ic.compute(g)
        rec = ic.export("prep", n_records=len(q))
    assert (rec["rx"] == rr.astype(int)).all() and (rec["ry"] == rr.astype(int)).all()
    ref32 = oracle.as_f32(c0.ref)
    for i in range(0, len(q), 5):
        r = int(rr[i])
        x, y = int(q["x"][i]), int(q["y"][i])
        sub = ref32[y - r:y + r + 1, x - r:x + r + 1].astype(np.float64)
        assert abs(float(rec["mean"][i]) - sub.mean()) <= 2 * np.spacing(np.float32(sub.mean()))


def test_12bit_left_alignment_matches_right(oracle, u16_scene):
    """A 12-bit image stored left-aligned (value x 16) and divided by 16 on
    the device gives the FFTCC, IC-GN and frame-driver results of the same
    image stored right-aligned, byte for byte (the division is exact)."""
    ref, tar = u16_scene
    assert ref.max() <= 4095 and tar.max() <= 4095
    refL, tarL = (ref << 4).astype(np.uint16), (tar << 4).astype(np.uint16)
    assert np.array_equal(oracle.as_f32(refL, "left"), oracle.as_f32(ref, "right"))
    q = fa.grid_queue(256, 256, 15, 9)
    out = {}
    for tag, (a, b, d) in {"right": (ref, tar, "right"), "left": (refL, tarL, "left")}.items():
        R, T = fa.Image2D(a, depth12=d), fa.Image2D(b, depth12=d)
        g = q.copy()
        with fa.FFTCC2D(15, 15) as f:
            f.setImages(R, T)
            f.compute(g)
        with engine(R, T) as ic:
            ic.compute(g)
            ref_f = ic.export("ref", (256, 256))
        with fa.FrameDriver(15) as drv:
            drv.set_reference(R)
            drv.set_pois(q)
            res, _ = drv.process(T)
        out[tag] = (g, ref_f, res)
    for k in range(3):
        assert out["left"][k].tobytes() == out["right"][k].tobytes(), k
    assert (out["left"][0]["result"][:, 2] > 0.9).mean() > 0.9


@pytest.mark.parametrize("depth,norm", [("native", "unit"), ("auto", "to16"), ("auto12", "none"), ("16", "to8"),
                                        ("left", "unit")])
def test_input_modes_convert_as_the_oracle(oracle, u16_scene, depth, norm):
    """The device's conversion of each input mode (loadImage2DMat,
    src/oc_image.cpp:125-217) equals the oracle's, bit for bit, for the
    reference and the target, and B-spline tables built from them too."""
    ref, tar = u16_scene
    if depth in ("auto12", "left"):
        ref, tar = (ref << 4).astype(np.uint16), (tar << 4).astype(np.uint16)
    h, w = ref.shape
    with engine(fa.Image2D(ref, depth, norm), fa.Image2D(tar, depth, norm)) as ic:
        got_r, got_t = ic.export("ref", (h, w)), ic.export("tar", (h, w))
        got_b = ic.export("bspline", (h, w))
    want_r, want_t = oracle.as_f32(ref, depth, norm), oracle.as_f32(tar, depth, norm)
    assert got_r.tobytes() == want_r.tobytes() and got_t.tobytes() == want_t.tobytes()
    assert got_b.tobytes() == oracle.bspline(want_t).tobytes()
