"""Output-integrity ledger (fasticgn_amd/ledger.py) on the host: the rules of
docs/output-integrity.md for publication, resume, repair, append, locking and
the read-only audit, on jobs whose payloads are deterministic bytes."""
import json
import os
from pathlib import Path

import numpy as np
import pytest

from fasticgn_amd import ledger
from fasticgn_amd.run import layout

MANIFEST = "manifest.opencorr-raw.json"


def make_inputs(tmp: Path, k: int):
    src = tmp / "inputs"
    src.mkdir(exist_ok=True)
    ref = src / "ref.npy"
    np.save(ref, np.arange(64, dtype=np.uint8).reshape(8, 8))
    tars = []
    for i in range(k):
        p = src / f"tar_{i:03d}.npy"
        np.save(p, np.full((8, 8), i, dtype=np.uint8))
        tars.append(p)
    return ref, tars


def contract_for(ref, tars, **opt):
    options = {"radius": 15, "step": 5, "max_iteration": 10, **opt}
    tl = [{"input": {"path": str(p.resolve()), **ledger.fingerprint(p)}, "output_name": p.name,
           "file": f"frame_{k:06d}.bin"} for k, p in enumerate(tars)]
    return ledger.make_contract({"library": "lib", "bytes": 1, "digest": "0" * 32},
                                {"path": str(ref.resolve()), **ledger.fingerprint(ref)}, options, tl)


def payload(k: int) -> bytes:
    return np.arange(36 * 10, dtype=np.uint8).tobytes() + bytes([k % 251]) * 36


def job(out: Path, contract, repair=False, stop=None, corrupt_replay=False):
    """Publish every frame in order (a replay from frame 0), then the manifest
    and the completion record; returns the session's counters."""
    outs, meta, _ = layout(contract, out)
    out.mkdir(parents=True, exist_ok=True)
    with ledger.OutputSession(out, contract, outs, meta, repair=repair) as s:
        s.begin()
        res = {"complete": s.complete, "verified": s.verified_files, "repaired": s.repaired_files,
               "archived": s.archived_residue, "kept": 0, "published": 0}
        if s.complete:
            return res
        for k, spec in enumerate(outs):
            if stop is not None and k >= stop:
                return res
            data = payload(k) if not corrupt_replay else payload(k + 1)
            res[s.publish(spec, data)] += 1
            s.publish_metadata(out / MANIFEST, json.dumps({"frames": k + 1}).encode())
        s.finish()
    return res


def audit(out):
    return ledger.audit_run_directory(out, layout)


def test_fresh_run_then_verified_complete(tmp_path):
    ref, tars = make_inputs(tmp_path, 4)
    c = contract_for(ref, tars)
    out = tmp_path / "out"
    r = job(out, c)
    assert r["published"] == 4 and not r["complete"]
    a = audit(out)
    assert a["status"] == "complete" and a["exit_code"] == 0 and a["frames_verified"] == 4
    assert not [p for p in out.iterdir() if p.name.startswith(ledger.TMP_PREFIX)]
    mtimes = {p.name: p.stat().st_mtime_ns for p in out.glob("frame_*.bin")}
    r = job(out, c)
    assert r["complete"] and r["verified"] == 5  # 4 payloads + the manifest
    assert {p.name: p.stat().st_mtime_ns for p in out.glob("frame_*.bin")} == mtimes


def test_interrupted_run_resumes_and_keeps_published_frames(tmp_path):
    ref, tars = make_inputs(tmp_path, 5)
    c = contract_for(ref, tars)
    out = tmp_path / "out"
    job(out, c, stop=2)
    a = audit(out)
    assert a["status"] == "incomplete" and a["missing_count"] == 3 and a["frames_verified"] == 2
    inode = (out / "frame_000000.bin").stat().st_ino
    r = job(out, c)
    assert r["kept"] == 2 and r["published"] == 3
    assert (out / "frame_000000.bin").stat().st_ino == inode
    assert audit(out)["status"] == "complete"
    fresh = tmp_path / "fresh"
    job(fresh, c)
    for k in range(5):
        assert (out / f"frame_{k:06d}.bin").read_bytes() == (fresh / f"frame_{k:06d}.bin").read_bytes()


def test_intent_without_payload_is_replayed(tmp_path):
    ref, tars = make_inputs(tmp_path, 3)
    c = contract_for(ref, tars)
    out = tmp_path / "out"
    job(out, c)
    os.unlink(out / "frame_000001.bin")  # a crash between intent and payload
    assert audit(out)["status"] == "incomplete"
    r = job(out, c)
    assert r["published"] == 1 and r["kept"] == 2
    assert audit(out)["status"] == "complete"


def test_damage_is_rejected_then_repaired(tmp_path):
    ref, tars = make_inputs(tmp_path, 3)
    c = contract_for(ref, tars)
    out = tmp_path / "out"
    job(out, c)
    p = out / "frame_000002.bin"
    b = bytearray(p.read_bytes())
    b[5] ^= 1  # same length, different content
    p.write_bytes(bytes(b))
    a = audit(out)
    assert a["status"] == "damaged" and a["mismatched"] == ["frame_000002.bin"]
    before = {q.name: q.read_bytes() for q in out.iterdir() if q.is_file()}
    with pytest.raises(ledger.IntegrityError, match="--repair"):
        job(out, c)
    assert {q.name: q.read_bytes() for q in out.iterdir() if q.is_file()} == before  # untouched
    r = job(out, c, repair=True)
    assert r["repaired"] == 1 and r["published"] == 1
    assert audit(out)["status"] == "complete"
    q = list((out / ledger.STATE_DIR / ledger.QUARANTINE).rglob("frame_000002.bin"))
    assert len(q) == 1 and q[0].read_bytes() == bytes(b)


def test_payload_without_intent_and_truncation(tmp_path):
    ref, tars = make_inputs(tmp_path, 3)
    c = contract_for(ref, tars)
    out = tmp_path / "out"
    job(out, c, stop=1)
    (out / "frame_000002.bin").write_bytes(payload(2))  # no intent
    p = out / "frame_000000.bin"
    p.write_bytes(p.read_bytes()[:-1])  # truncated
    a = audit(out)
    assert a["status"] == "damaged" and a["no_intent"] == ["frame_000002.bin"]
    assert a["mismatched"] == ["frame_000000.bin"]
    with pytest.raises(ledger.IntegrityError):
        job(out, c)
    job(out, c, repair=True)
    assert audit(out)["status"] == "complete"


def test_changed_options_and_inputs_are_rejected(tmp_path):
    ref, tars = make_inputs(tmp_path, 3)
    out = tmp_path / "out"
    job(out, contract_for(ref, tars))
    with pytest.raises(ledger.IntegrityError, match=r"options\.max_iteration: 10 -> 11"):
        job(out, contract_for(ref, tars, max_iteration=11))
    with pytest.raises(ledger.IntegrityError, match="target 1 of the recorded 3 differs"):
        job(out, contract_for(ref, [tars[0], tars[2], tars[1]]))
    with pytest.raises(ledger.IntegrityError, match="a shorter list is not a resume"):
        job(out, contract_for(ref, tars[:2]))
    # an input rewritten after the run: outputs intact, inputs changed
    np.save(tars[1], np.full((8, 8), 99, dtype=np.uint8))
    a = audit(out)
    assert a["status"] == "inputs_changed" and a["inputs"]["changed"] == [str(tars[1].resolve())]


def test_append_extends_the_contract(tmp_path):
    ref, tars = make_inputs(tmp_path, 5)
    out = tmp_path / "out"
    job(out, contract_for(ref, tars[:3]))
    long_c = contract_for(ref, tars)
    # cut off right after the first appended payload (before promotion)
    outs, meta, _ = layout(long_c, out)
    with ledger.OutputSession(out, long_c, outs, meta) as s:
        s.begin()
        assert (out / ledger.STATE_DIR / ledger.PENDING).exists()
        for k in range(3):
            assert s.publish(outs[k], payload(k)) == "kept"
        # publish frame 3 without the promotion that follows it
        s._pending = False
        s.publish(outs[3], payload(3))
    a = audit(out)
    assert a["status"] == "incomplete" and a["appended_published"] == ["frame_000003.bin"]
    with pytest.raises(ledger.IntegrityError, match="untracked entry"):
        job(out, contract_for(ref, tars[:3]))  # the short list no longer accounts for frame 3
    r = job(out, long_c)
    assert r["kept"] == 4 and r["published"] == 1
    a = audit(out)
    assert a["status"] == "complete" and a["frames_expected"] == 5
    assert not (out / ledger.STATE_DIR / ledger.PENDING).exists()


def test_untracked_residue_lock_and_statuses(tmp_path):
    ref, tars = make_inputs(tmp_path, 2)
    c = contract_for(ref, tars)
    out = tmp_path / "out"
    job(out, c)
    (out / "notes.bin").write_bytes(b"x")
    a = audit(out)
    assert a["status"] == "damaged" and a["untracked_outputs"] == ["notes.bin"]
    with pytest.raises(ledger.IntegrityError, match="untracked entry in output directory: notes.bin"):
        job(out, c)
    os.unlink(out / "notes.bin")
    (out / ".opencorr-tmp-123-4").write_bytes(b"partial")
    a = audit(out)
    assert a["status"] == "complete" and a["residue"] == [".opencorr-tmp-123-4"]
    r = job(out, c)
    assert r["complete"] and r["archived"] == 1 and not (out / ".opencorr-tmp-123-4").exists()
    # a second session while one holds the lock
    outs, meta, _ = layout(c, out)
    with ledger.OutputSession(out, c, outs, meta) as s:
        s.begin()
        assert audit(out)["status"] == "locked"
        with pytest.raises(ledger.IntegrityError, match="already locked"):
            ledger.OutputSession(out, c, outs, meta).begin()
    assert audit(out)["status"] == "complete"
    assert audit(tmp_path / "nope")["status"] == "not_found"
    assert audit(tmp_path / "inputs")["status"] == "not_a_run"


def test_replay_with_different_bytes_is_an_error(tmp_path):
    ref, tars = make_inputs(tmp_path, 3)
    c = contract_for(ref, tars)
    out = tmp_path / "out"
    job(out, c, stop=2)
    with pytest.raises(ledger.IntegrityError, match="differs from the recorded payload"):
        job(out, c, corrupt_replay=True)


def test_pre_ledger_outputs_and_contract_tampering(tmp_path):
    ref, tars = make_inputs(tmp_path, 2)
    c = contract_for(ref, tars)
    out = tmp_path / "out"
    out.mkdir()
    (out / "frame_000000.bin").write_bytes(payload(0))
    with pytest.raises(ledger.IntegrityError, match="pre-ledger"):
        job(out, c)
    out2 = tmp_path / "out2"
    job(out2, c)
    cp = out2 / ledger.STATE_DIR / ledger.CONTRACT
    doc = json.loads(cp.read_text())
    doc["options"]["radius"] = 16  # edited without its digest
    cp.write_text(json.dumps(doc))
    with pytest.raises(ledger.IntegrityError, match="digest mismatch"):
        job(out2, c)
    assert audit(out2)["status"] == "damaged"


def test_atomic_write_never_replaces_a_payload(tmp_path):
    p = tmp_path / "a.bin"
    ledger.atomic_write(p, b"one", replace=False)
    with pytest.raises(FileExistsError):
        ledger.atomic_write(p, b"two", replace=False)
    assert p.read_bytes() == b"one"
    ledger.atomic_write(p, b"three", replace=True, durability="fsync")
    assert p.read_bytes() == b"three"
    assert [q.name for q in tmp_path.iterdir()] == ["a.bin"]


def test_describe_json_differences_limits():
    a = {"o": {str(i): i for i in range(12)}}
    b = {"o": {str(i): i + 1 for i in range(12)}}
    s = ledger.describe_json_differences(a, b)
    assert s.count("->") == 8 and s.endswith("and 4 more")
