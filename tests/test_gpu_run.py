"""The multi-frame runner with the output-integrity ledger on the device
(python -m fasticgn_amd.run): a 12-bit PGM sequence run in one go, and the
same command stopped after three frames and run again, give byte-identical
frame files (the resume replays from the first frame and keeps the three
published payloads); the frames equal the frame driver's own records; the
audit reports complete; rerunning a complete job computes nothing."""
import json

import numpy as np
import pytest

import fasticgn_amd as fa
from fasticgn_amd import rawio, run, synth
from fasticgn_amd.opencorr import grid_queue, write_pgm

pytestmark = pytest.mark.gpu

FRAMES = (0, 30, 60, 90, 120, 150)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def sequence(tmp_path_factory):
    d = tmp_path_factory.mktemp("seq")
    seq = synth.config2(size=256, frames=500)
    write_pgm(d / "ref.pgm", seq.ref)
    names = []
    for k in FRAMES:
        name = f"tar_{k:03d}.pgm"
        write_pgm(d / name, seq.frame(k))
        names.append(name)
    (d / "targets.txt").write_text("\n".join(names) + "\n")
    return d, seq


def cmd(d, out, *extra):
    return ["--ref", str(d / "ref.pgm"), "--tar-list", str(d / "targets.txt"), "--output-dir", str(out),
            "--radius", "15", "--step", "7", "--input-depth", "12", "--input-align", "right", *extra]


def frames(out):
    return [(out / f"frame_{k:06d}.bin").read_bytes() for k in range(len(FRAMES))]


def test_runner_resume_is_byte_identical(sequence, tmp_path, capsys):
    d, seq = sequence
    full, part = tmp_path / "full", tmp_path / "part"
    assert run.main(cmd(d, full)) == 0
    st = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    say("full run", st)
    assert st["published"] == len(FRAMES)
    assert run.main(cmd(d, part, "--stop-after", "3")) == 0
    st = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert st["status"] == "stopped" and st["published"] == 3
    assert run.main(["--verify-only", str(part)]) == 1
    rep = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert rep["status"] == "incomplete" and rep["missing_count"] == 3
    assert run.main(cmd(d, part)) == 0
    st = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    say("resumed run", st)
    assert st["kept"] == 3 and st["published"] == 3
    assert frames(part) == frames(full)
    assert run.main(["--verify-only", str(part)]) == 0
    rep = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert rep["status"] == "complete" and rep["frames_verified"] == len(FRAMES)
    assert run.main(cmd(d, part)) == 0
    st = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert st["status"] == "complete"
    # the frames are the frame driver's records of the same sequence
    with fa.FrameDriver(15) as drv:
        ref = fa.Image2D(seq.ref, depth12="right")
        drv.set_reference(ref)
        drv.set_pois(grid_queue(256, 256, 15, 7))
        for k, b in zip(FRAMES, frames(full)):
            drv.process(fa.Image2D(seq.frame(k), depth12="right"))
            assert drv.records().tobytes() == b
    man = json.loads((full / rawio.RAW_MANIFEST_NAME).read_text())
    assert [f["file"] for f in man["frames"]] == [f"frame_{k:06d}.bin" for k in range(len(FRAMES))]
    assert all(f["rows"] == len(grid_queue(256, 256, 15, 7)) for f in man["frames"])
    # a changed option is refused and the directory is left as it was
    assert run.main(cmd(d, part, "--max-iteration", "11")) == 1
    out = capsys.readouterr().out
    assert "options.max_iteration: 10 -> 11" in out
    assert frames(part) == frames(full)
    assert np.frombuffer(frames(full)[0], dtype=rawio.RAW_DTYPE)["zncc"].max() > 0.9
