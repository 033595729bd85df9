"""The OpenCorr-shaped C++ header compiles against the C ABI and links with
libfasticgn.so (CPU); the probe runs on the device (GPU)."""
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
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
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
