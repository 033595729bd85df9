"""Raw output tree (host only): frame files through fi_write_raw and the
opencorr-gpu-raw-v1 manifest, read back with the checks of the reference's
reader (pipeline/fasticgn_io.py:132-145 read_raw_frame, :256-268
read_manifest, :274-292 read_frame_from_manifest) restated here."""
import json

import numpy as np
import pytest

from fasticgn_amd import rawio
from fasticgn_amd.opencorr import poi_queue


def sample_queue(n=37, seed=0):
    rng = np.random.default_rng(seed)
    q = poi_queue(np.stack([rng.integers(20, 500, n), rng.integers(20, 500, n)], axis=1).astype(np.float32))
    d = q["deformation"]
    d[:, 0] = rng.normal(0, 3, n)
    d[:, 6] = rng.normal(0, 3, n)
    d[:, 1], d[:, 2], d[:, 7], d[:, 8] = (rng.normal(0, 1e-3, n) for _ in range(4))
    q["result"][:, 2] = rng.uniform(-5, 1, n)
    q["result"][3, 2] = np.nan
    return q


def read_manifest_like_reference(path):
    m = json.loads(path.read_text(encoding="utf-8"))
    assert isinstance(m, dict)
    assert m.get("format") == "opencorr-gpu-raw-v1"
    assert m.get("endianness") == "little"
    assert m.get("record_bytes") == 36
    assert isinstance(m.get("frames"), list)
    return m


def test_record_layout_by_hand(tmp_path):
    q = sample_queue()
    n = rawio.write_raw(q, tmp_path / "f.bin")
    raw = (tmp_path / "f.bin").read_bytes()
    assert n == len(raw) == 36 * len(q)
    for i in (0, 3, len(q) - 1):
        rec = raw[36 * i: 36 * (i + 1)]
        assert np.frombuffer(rec[:8], "<i4").tolist() == [int(q["x"][i]), int(q["y"][i])]
        f = np.frombuffer(rec[8:], "<f4")
        d = q["deformation"][i]
        want = np.array([d[0], d[6], d[1], d[2], d[7], d[8], q["result"][i, 2]], dtype=np.float32)
        assert f.tobytes() == want.tobytes()


def test_tree_reads_back(tmp_path):
    frames = [sample_queue(seed=s) for s in range(3)]
    qa = [{"bad_ratio": 0.01, "baseline_bad_ratio": float("nan"), "fallback_threshold": float("nan"),
           "fallback_used": 0, "skip_fftcc_used": 0, "valid_ratio": 0.9, "iteration_mean": 3.1,
           "iteration_max": 7, "poi_fallback_count": 0, "poi_fallback_recovered": 0},
          {"bad_ratio": 0.02, "baseline_bad_ratio": 0.01, "fallback_threshold": 0.04, "fallback_used": 0,
           "skip_fftcc_used": 1, "valid_ratio": 0.91, "iteration_mean": 2.0, "iteration_max": 5,
           "poi_fallback_count": 4, "poi_fallback_recovered": 3},
          None]
    with rawio.RawWriter(tmp_path, ref_path="ref.tif", subset_radius=15, grid_space=5) as w:
        for k, (q, s) in enumerate(zip(frames, qa)):
            w.write_frame(q, s, tar_path=f"tar_{k}.tif")
    m = read_manifest_like_reference(tmp_path / rawio.RAW_MANIFEST_NAME)
    assert [f["file"] for f in m["frames"]] == [f"frame_{k:06d}.bin" for k in range(3)]
    assert m["fields"][0] == {"name": "x", "dtype": "int32", "offset": 0}
    assert m["fields"][-1] == {"name": "zncc", "dtype": "float32", "offset": 32}
    assert m["frames"][0]["baseline_bad_ratio"] is None and m["frames"][1]["baseline_bad_ratio"] == 0.01
    assert m["frames"][1]["skip_fftcc_used"] is True and m["frames"][1]["poi_fallback_count"] == 4
    assert m["frames"][2]["bad_ratio"] is None
    for entry, q in zip(m["frames"], frames):
        rec = rawio.read_raw_frame(tmp_path / entry["file"])
        assert rec.shape[0] == entry["rows"] and rec.nbytes == entry["bytes"]
        assert np.array_equal(rec["x"], q["x"].astype(np.int32))
        assert np.array_equal(rec["u"], q["deformation"][:, 0], equal_nan=True)
        assert np.array_equal(rec["vy"], q["deformation"][:, 8], equal_nan=True)
        assert np.array_equal(rec["zncc"], q["result"][:, 2], equal_nan=True)


def test_truncated_frame_is_rejected(tmp_path):
    rawio.write_raw(sample_queue(), tmp_path / "frame_000000.bin")
    p = tmp_path / "frame_000000.bin"
    p.write_bytes(p.read_bytes()[:-5])
    with pytest.raises(ValueError):
        rawio.read_raw_frame(p)


def test_empty_frame(tmp_path):
    assert rawio.write_raw(poi_queue(np.zeros((0, 2), np.float32)), tmp_path / "e.bin") == 0
    assert (tmp_path / "e.bin").read_bytes() == b""
