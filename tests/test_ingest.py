"""Image input modes (loadImage2DMat, src/oc_image.cpp:125-217) on the CPU:
the oracle's conversion pinned by closed forms, the library's mode
resolution (fi_image_resolve, no device work) against it, and the PGM files
Image2D.load reads."""
import numpy as np
import pytest

import fasticgn_amd as fa
from fasticgn_amd.opencorr import read_pgm, write_pgm


@pytest.fixture
def u12():
    rng = np.random.default_rng(9)
    return rng.integers(0, 4096, (37, 53), dtype=np.uint16)


def test_oracle_closed_forms(oracle, u12):
    u8 = (u12 >> 4).astype(np.uint8)
    # 8-bit as stored, and normalised to [0, 1]: float(v) * float(1 / 255)
    assert np.array_equal(oracle.as_f32(u8), u8.astype(np.float32))
    unit = oracle.as_f32(u8, normalize="unit")
    assert np.array_equal(unit, u8.astype(np.float32) * np.float32(1 / 255))
    # two roundings (the scale, the product): within two units in the last place
    assert (np.abs(unit.astype(np.float64) - u8 / 255.0) <= 2 * np.spacing(unit)).all()
    # 16-bit data up to 4095: 'auto' infers 12-bit right-aligned
    assert np.array_equal(oracle.as_f32(u12, "auto", "to16"), u12.astype(np.float32) * np.float32(65535 / 4095))
    assert np.array_equal(oracle.as_f32(u12, "auto"), u12.astype(np.float32))
    # above 4095: 'auto' is 16-bit; 'auto12' takes the data as left-aligned
    left = (u12 << 4).astype(np.uint16)
    left[0, 0] = 65535 - 15
    assert np.array_equal(oracle.as_f32(left, "auto", "unit"), left.astype(np.float32) * np.float32(1 / 65535))
    assert np.array_equal(oracle.as_f32(left, "auto12"), left.astype(np.float32) / 16)
    assert np.array_equal(oracle.as_f32(left, "16", "to8"), left.astype(np.float32) * np.float32(255 / 65535))
    with pytest.raises(ValueError, match="above 4095"):
        oracle.as_f32(left, "right")
    # 12-bit modes on 8-bit data (right-aligned: as stored; left: / 16)
    assert np.array_equal(oracle.as_f32(u8, "left"), u8.astype(np.float32) / 16)
    assert np.array_equal(oracle.as_f32(u8, "right", "unit"), u8.astype(np.float32) * np.float32(1 / 4095))
    with pytest.raises(ValueError):
        oracle.as_f32(u8.astype(np.float32), "auto")


CASES = [("native", "none"), ("native", "unit"), ("right", "to8"), ("left", "none"), ("auto12", "to16"),
         ("auto", "none"), ("auto", "unit"), ("16", "to8")]


@pytest.mark.parametrize("depth,norm", CASES)
@pytest.mark.parametrize("kind", ["u8", "u12", "u16"])
def test_resolution_matches_oracle(oracle, u12, kind, depth, norm):
    px = {"u8": (u12 >> 4).astype(np.uint8), "u12": u12, "u16": (u12 << 4).astype(np.uint16) | 15}[kind]
    try:
        want = oracle.as_f32(px, depth, norm)
    except ValueError:
        with pytest.raises(ValueError):
            fa.Image2D(px, depth, norm).info()
        return
    img = fa.Image2D(px, depth, norm)
    info = img.info()
    assert img.as_float().tobytes() == want.tobytes()
    assert info["effective_bits"] in (8, 12, 16)


def test_pgm_round_trip(tmp_path, u12):
    for px in (u12, (u12 >> 4).astype(np.uint8)):
        write_pgm(tmp_path / "a.pgm", px)
        back = read_pgm(tmp_path / "a.pgm")
        assert back.dtype == px.dtype and np.array_equal(back, px)
    (tmp_path / "b.pgm").write_bytes(b"P2\n# comment\n3 2\n4095\n1 2 3\n4095 0 7\n")
    assert np.array_equal(read_pgm(tmp_path / "b.pgm"), np.array([[1, 2, 3], [4095, 0, 7]], np.uint16))
    (tmp_path / "c.pgm").write_bytes(b"P5\n3 2\n255\n\x01\x02")
    with pytest.raises(ValueError, match="Fail to load file"):
        read_pgm(tmp_path / "c.pgm")


def test_image_load_options(tmp_path, u12):
    write_pgm(tmp_path / "d.pgm", u12)
    default = fa.Image2D.load(tmp_path / "d.pgm")  # 8-bit: the high byte
    assert default.data.dtype == np.uint8 and np.array_equal(default.data, (u12 >> 8).astype(np.uint8))
    auto = fa.Image2D.load(tmp_path / "d.pgm", depth="auto", normalize="unit")
    info = auto.info()
    assert info["effective_bits"] == 12 and not info["left_aligned"] and info["scale"] == 1 / 4095
    left = fa.Image2D.load(tmp_path / "d.pgm", depth="12", alignment="left")
    assert left.info()["left_aligned"] == 1
