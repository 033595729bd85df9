"""The parity gate itself (tests/parity.py), on hand-built frames.

The first thirteen cases restate the reference's own pins of its gate
(smoke/test_check_cpu_gpu_parity.py:17-145): each plants one fault in an
otherwise identical frame and expects the verdict the reference expects.  The
rest pin what this project adds: the gradient fields, validity identity and
the iteration-count bound of compare().
"""
import copy
import json
import math
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from fasticgn_amd.opencorr import poi_queue
from tests.parity import compare, compare_frame, rows_from_pois

CHECKER = Path(__file__).with_name("parity.py")


@pytest.fixture
def cpu():
    return {(x, 10): {"u": 5.0, "v": 0.0, "zncc": 0.99} for x in range(200)}


# --------------------------------------------------- the reference's cases --

def test_matching_frame_passes(cpu):
    assert compare_frame(cpu, cpu)["pass"]


def test_entire_lost_frame_fails_even_after_a_good_frame(cpu):
    good = compare_frame(cpu, cpu)
    gpu = copy.deepcopy(cpu)
    for row in gpu.values():
        row["zncc"] = -4.0
    bad = compare_frame(cpu, gpu)
    assert good["pass"] and not bad["pass"]
    assert bad["cpu_valid_gpu_invalid"] == 200


def test_lost_ten_percent_cannot_hide_in_zero_displacement_error(cpu):
    gpu = copy.deepcopy(cpu)
    for key in list(gpu)[:20]:
        gpu[key]["zncc"] = -4.0
    res = compare_frame(cpu, gpu)
    assert res["median_abs_du"] == 0.0
    assert not res["pass"]


def test_tail_error_cannot_hide_in_median(cpu):
    gpu = copy.deepcopy(cpu)
    for key in list(gpu)[:4]:
        gpu[key]["u"] += 1.0
    res = compare_frame(cpu, gpu)
    assert res["median_abs_du"] == 0.0
    assert not res["pass"]


def test_nonfinite_displacement_is_invalid_even_with_positive_zncc(cpu):
    gpu = copy.deepcopy(cpu)
    for row in gpu.values():
        row["u"] = float("nan")
    assert not compare_frame(cpu, gpu)["pass"]


def test_all_invalid_on_both_sides_is_not_parity_evidence(cpu):
    rows = copy.deepcopy(cpu)
    for row in rows.values():
        row["zncc"] = -3.0
    assert not compare_frame(rows, rows)["pass"]


def test_empty_common_domain_fails(cpu):
    assert not compare_frame(cpu, {})["pass"]


def test_single_outlier_fails_max_when_p99_is_zero(cpu):
    gpu = copy.deepcopy(cpu)
    gpu[(0, 10)]["u"] += 1.0
    res = compare_frame(cpu, gpu)
    assert res["p99_abs_du"] == 0.0
    assert not res["pass"]
    assert compare_frame(cpu, gpu, max_threshold=1.0)["pass"]


def write_frames(folder: Path, names, u=5.0):
    folder.mkdir(parents=True, exist_ok=True)
    for name in names:
        rows = "".join(f"{x},10,{u},0,0.99\n" for x in range(200))
        (folder / name).write_text("x,y,u,v,zncc\n" + rows)


def run_checker(cpu_dir: Path, gpu_dir: Path, *extra):
    rep = cpu_dir.parent / "report.json"
    run = subprocess.run([sys.executable, str(CHECKER), "--cpu-dir", str(cpu_dir), "--gpu-dir", str(gpu_dir),
                          "--json-report", str(rep), *extra], capture_output=True, text=True)
    assert "Traceback" not in run.stderr, run.stderr
    return run.returncode, json.loads(rep.read_text())


def test_default_pairing_still_uses_zero_based_names(tmp_path):
    write_frames(tmp_path / "cpu", ["frame_000000.csv", "frame_000001.csv"])
    write_frames(tmp_path / "gpu", ["frame_000000.csv", "frame_000001.csv"])
    code, rep = run_checker(tmp_path / "cpu", tmp_path / "gpu", "--frames", "2")
    assert code == 0
    assert [f["frame"] for f in rep["frames"]] == ["frame_000000.csv", "frame_000001.csv"]
    assert rep["pairing"] == "index"


def test_glob_pairs_arbitrary_frame_names(tmp_path):
    names = ["frame_072425.csv", "frame_072424.csv", "frame_072426.csv"]
    write_frames(tmp_path / "cpu", names)
    write_frames(tmp_path / "gpu", names)
    (tmp_path / "gpu" / "notes.txt").write_text("not a frame")
    code, rep = run_checker(tmp_path / "cpu", tmp_path / "gpu", "--glob", "frame_*.csv")
    assert code == 0
    assert [f["frame"] for f in rep["frames"]] == sorted(names)
    code, rep = run_checker(tmp_path / "cpu", tmp_path / "gpu", "--glob", "frame_*.csv", "--frames", "2")
    assert code == 0 and len(rep["frames"]) == 2


def test_glob_fails_a_name_present_on_one_side_only(tmp_path):
    write_frames(tmp_path / "cpu", ["frame_000002.csv", "frame_000003.csv"])
    write_frames(tmp_path / "gpu", ["frame_000002.csv"])
    code, rep = run_checker(tmp_path / "cpu", tmp_path / "gpu", "--glob", "frame_*.csv")
    assert code == 1
    assert {f["frame"]: f["pass"] for f in rep["frames"]} == {"frame_000002.csv": True,
                                                               "frame_000003.csv": False}


def test_glob_without_any_match_fails(tmp_path):
    (tmp_path / "cpu").mkdir()
    (tmp_path / "gpu").mkdir()
    code, rep = run_checker(tmp_path / "cpu", tmp_path / "gpu", "--glob", "frame_*.csv")
    assert code == 1
    assert rep["frames"] == [] and not rep["pass"]


def test_missing_csv_column_is_reported_without_traceback(tmp_path):
    (tmp_path / "frame_000000.csv").write_text("x,y,u,v,zncc\n1,2,3\n")
    rep = tmp_path / "report.json"
    run = subprocess.run([sys.executable, str(CHECKER), "--cpu-dir", str(tmp_path), "--gpu-dir", str(tmp_path),
                          "--frames", "1", "--json-report", str(rep)], capture_output=True, text=True)
    assert run.returncode == 1
    assert "Traceback" not in run.stderr
    assert not json.loads(rep.read_text())["frames"][0]["pass"]


# ------------------------------------------------- what the project adds --

@pytest.fixture
def pois():
    q = poi_queue([(x, y) for y in range(20, 220, 10) for x in range(20, 220, 10)])
    rng = np.random.default_rng(3)
    q["deformation"][:, [0, 6]] = rng.uniform(-3, 3, (len(q), 2))
    q["deformation"][:, [1, 2, 7, 8]] = rng.uniform(-0.01, 0.01, (len(q), 4))
    q["result"][:, 2] = 0.95
    q["result"][:, 3] = 3
    q["result"][:, 4] = 2e-4
    return q


@pytest.mark.parametrize("col,name", [(1, "ux"), (2, "uy"), (7, "vx"), (8, "vy")])
def test_gradient_tail_error_fails(pois, col, name):
    """A 0.02 error in one gradient on 5 % of the POIs passes the u/v-only
    gate and fails the project's gate, by the p99 rule of that field."""
    g = pois.copy()
    g["deformation"][::20, col] += 0.02
    assert compare_frame(rows_from_pois(pois), rows_from_pois(g))["pass"]
    res = compare(pois, g)
    assert not res["pass"]
    assert res[f"p99_d{name}"] == pytest.approx(0.02, rel=1e-3)
    assert any("p99" in e for e in res["errors"])


def test_gradient_nan_on_a_valid_poi_fails(pois):
    g = pois.copy()
    g["deformation"][7, 2] = np.nan
    assert not compare(pois, g)["pass"]


def test_status_identity_is_stricter_than_retention(pois):
    """-4 against -5 on one POI: both invalid, so the reference's rules pass,
    the identity of the flags does not."""
    c, g = pois.copy(), pois.copy()
    c["result"][5, 2], g["result"][5, 2] = -4.0, -5.0
    res = compare(c, g)
    assert res["status_mismatch"] == 1 and not res["pass"]
    assert "validity flags differ" in res["errors"]


def test_iteration_mismatch_is_reported_and_bounded(pois):
    g = pois.copy()
    g["result"][:10, 3] += 1
    res = compare(pois, g)
    assert res["pass"] and res["iteration_mismatch"] == 10 and res["max_diteration"] == 1.0
    assert res["iteration_mismatch_frac"] == pytest.approx(10 / len(pois))
    assert not compare(pois, g, max_iteration_mismatch=0.01)["pass"]
    assert compare(pois, g, max_iteration_mismatch=0.05)["pass"]


def test_nearest_rank_is_not_interpolated():
    from tests.parity import nearest_rank
    vals = list(range(1, 101))
    assert nearest_rank(vals, 0.99) == 99
    assert nearest_rank([0.0] * 199 + [1.0], 0.99) == 0.0
    assert nearest_rank([5.0], 0.99) == 5.0
    assert math.isclose(nearest_rank([3.0, 1.0, 2.0], 0.5), 2.0)


def test_duplicate_location_is_an_error(pois):
    q = pois.copy()
    q["x"][1], q["y"][1] = q["x"][0], q["y"][0]
    with pytest.raises(ValueError, match="duplicate"):
        compare(q, q)
