"""The OpenCorr-shaped C++ header compiles against the C ABI and links with
libfasticgn.so (CPU); the probes run on the device (GPU).  Where it was
built (oracle/ref_probe.mk, from the reference checkout), the reference's own
smoke/library_boundary_probe.cpp, compiled unchanged against these headers,
runs too."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "cpp" / "library_probe.cpp"
LIBDIR = ROOT / "fasticgn_amd"


def build(tmp_path: Path) -> Path:
    exe = tmp_path / "library_probe"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", f"-I{ROOT / 'include'}", str(SRC),
                    f"-L{LIBDIR}", "-lfasticgn", f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)], check=True)
    return exe


def test_cpp_header_compiles_and_links(tmp_path):
    assert build(tmp_path).exists()


@pytest.mark.gpu
def test_cpp_library_probe_runs(tmp_path):
    exe = build(tmp_path)
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(out.stdout, flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FAIL" not in out.stdout


FRAME_SRC = ROOT / "tests" / "cpp" / "frame_probe.c"


def build_frame_probe(tmp_path: Path) -> Path:
    exe = tmp_path / "frame_probe"
    subprocess.run(["gcc", "-std=c11", "-O1", "-Wall", f"-I{ROOT / 'include'}", str(FRAME_SRC),
                    f"-L{LIBDIR}", "-lfasticgn", f"-Wl,-rpath,{LIBDIR}", "-lm", "-o", str(exe)], check=True)
    return exe


def test_c_frame_probe_compiles_and_links(tmp_path):
    assert build_frame_probe(tmp_path).exists()


@pytest.mark.gpu
def test_c_frame_probe_runs(tmp_path):
    exe = build_frame_probe(tmp_path)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(out.stdout, flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FAIL" not in out.stdout


REF_PROBE = ROOT / "oracle" / "_ref" / "library_boundary_probe"


@pytest.mark.gpu
def test_reference_library_probe_runs_unchanged():
    """The reference's FFTCC2D lifecycle probe (radius and thread-count
    guards, cached layout, the LCG known answer, DIC::setSubset rejected)
    against this library."""
    if not REF_PROBE.exists():
        pytest.skip("the reference's probe was not built (no reference checkout where build() ran)")
    out = subprocess.run([str(REF_PROBE)], capture_output=True, text=True, timeout=120)
    print(out.stdout, flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FAIL" not in out.stdout and "library boundary failures=0" in out.stdout
