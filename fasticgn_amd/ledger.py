"""Output integrity for multi-frame runs: atomic publication, an intent ledger,
resume and a read-only audit.

Restates the contract of the reference's CLI I/O library
(src/cli_output_integrity.h:23-137; docs/output-integrity.md) for this
package's raw output tree:

- every payload and manifest is written to a private temporary file
  ``.opencorr-tmp-<pid>-<n>`` in its destination directory (O_EXCL) and
  published by ``link`` (payloads: never replace an existing file) or
  ``rename`` (metadata: atomic replace); ``durability="fsync"`` adds an fsync
  of the file and its directory;
- a per-file intent record (key, rows, bytes, XXH3-128 digest) is published in
  ``.opencorr-state/intents/`` before its payload, so an intent without a file
  is a frame to replay and a file without an intent is damage;
- ``.opencorr-state/contract.json`` records what decides the output bytes
  (library fingerprint, reference, options, target list); a resume with a
  different contract is rejected naming the fields that differ, and a target
  list may only grow by appending (``contract.pending.json`` until the first
  appended payload is published);
- a resume re-verifies every recorded payload by size and digest; a complete
  matching run computes nothing (``complete``); an incomplete one replays from
  the first frame and keeps the verified payloads untouched;
- ``repair=True`` moves damaged payloads to ``.opencorr-state/quarantine/``
  and republishes them; crash residue is archived there on every resume;
- ``.opencorr-state/complete.json`` is published last, with the contract's
  digest and the manifests' fingerprints;
- ``audit_run_directory`` (``--verify-only``) judges a directory without
  changing it: complete, incomplete, damaged, inputs_changed, locked,
  not_a_run, not_found.

Scope: POSIX local filesystems (flock, link, rename, directory fsync).  The
digests detect accidental corruption; they are not authentication.
"""
from __future__ import annotations

import errno
import fcntl
import json
import os
import re
import time
from dataclasses import dataclass
from pathlib import Path

import xxhash

STATE_DIR = ".opencorr-state"
TMP_PREFIX = ".opencorr-tmp-"
CONTRACT = "contract.json"
PENDING = "contract.pending.json"
COMPLETE = "complete.json"
INTENTS = "intents"
QUARANTINE = "quarantine"
LOCK = "lock"
CONTRACT_SCHEMA = "fasticgn.contract/1"
INTENT_SCHEMA = "fasticgn.intent/1"
COMPLETE_SCHEMA = "fasticgn.complete/1"
AUDIT_SCHEMA = "fasticgn.verify-only/1"
LIST_LIMIT = 20
_RESIDUE = re.compile(r"^\.opencorr-tmp-(\d+)-(\d+)$")
_HEX32 = re.compile(r"^[0-9a-f]{32}$")


class IntegrityError(RuntimeError):
    """A resume or publication the ledger refuses (the directory is unchanged)."""


# ---------------------------------------------------------------------------
# digests and atomic files
# ---------------------------------------------------------------------------
def digest(data: bytes) -> str:
    """XXH3-128 of the bytes, 32 lowercase hex digits."""
    return xxhash.xxh3_128(data).hexdigest()


def fingerprint(path) -> dict:
    """{"bytes", "digest"} of a regular file's content."""
    p = Path(path)
    h = xxhash.xxh3_128()
    n = 0
    with open(p, "rb") as f:
        while True:
            b = f.read(1 << 22)
            if not b:
                break
            h.update(b)
            n += len(b)
    return {"bytes": n, "digest": h.hexdigest()}


def canonical_json(value) -> bytes:
    return (json.dumps(value, sort_keys=True, separators=(",", ":"), ensure_ascii=False) + "\n").encode()


_counter = [0]


def _fsync_dir(d: Path) -> None:
    fd = os.open(d, os.O_RDONLY | os.O_DIRECTORY)
    try:
        os.fsync(fd)
    finally:
        os.close(fd)


def atomic_write(path, data: bytes, replace: bool = True, durability: str = "rename") -> None:
    """Publish data at path through a private temporary file in the same
    directory: rename over the destination (replace) or hard link to it
    (no replace: FileExistsError if it exists).  A reader never sees a
    partial file."""
    path = Path(path)
    d = path.parent
    while True:
        _counter[0] += 1
        tmp = d / f"{TMP_PREFIX}{os.getpid()}-{_counter[0]}"
        try:
            fd = os.open(tmp, os.O_WRONLY | os.O_CREAT | os.O_EXCL, 0o644)
            break
        except FileExistsError:
            continue
    try:
        view = memoryview(data)
        while view:
            k = os.write(fd, view)
            view = view[k:]
        if durability == "fsync":
            os.fsync(fd)
    finally:
        os.close(fd)
    try:
        if replace:
            os.replace(tmp, path)
        else:
            os.link(tmp, path)
            os.unlink(tmp)
    except BaseException:
        if tmp.exists():
            os.unlink(tmp)
        raise
    if durability == "fsync":
        _fsync_dir(d)


def read_json(path) -> dict:
    """Strict JSON object (duplicate keys rejected)."""
    def pairs(kv):
        out = {}
        for k, v in kv:
            if k in out:
                raise ValueError(f"duplicate key {k!r}")
            out[k] = v
        return out
    v = json.loads(Path(path).read_text(encoding="utf-8"), object_pairs_hook=pairs)
    if not isinstance(v, dict):
        raise ValueError("not a JSON object")
    return v


def describe_json_differences(recorded, requested, limit: int = 8) -> str:
    """"a.b: old -> new; ..." for the leaves that differ, at most limit."""
    diffs: list[str] = []

    def walk(a, b, prefix):
        if isinstance(a, dict) and isinstance(b, dict):
            for k in sorted(set(a) | set(b)):
                walk(a.get(k, "<absent>"), b.get(k, "<absent>"), f"{prefix}.{k}" if prefix else k)
        elif isinstance(a, list) and isinstance(b, list) and len(a) == len(b):
            for i, (x, y) in enumerate(zip(a, b)):
                walk(x, y, f"{prefix}[{i}]")
        elif a != b:
            diffs.append(f"{prefix}: {json.dumps(a)} -> {json.dumps(b)}")

    walk(recorded, requested, "")
    out = "; ".join(diffs[:limit])
    if len(diffs) > limit:
        out += f"; and {len(diffs) - limit} more"
    return out


# ---------------------------------------------------------------------------
# contract
# ---------------------------------------------------------------------------
def make_contract(binary: dict, reference: dict, options: dict, targets: list[dict]) -> dict:
    """targets: [{"input": {"path", "bytes", "digest"}, "output_name", "file"}]"""
    body = {"schema": CONTRACT_SCHEMA, "binary": binary, "reference": reference, "options": options,
            "targets": targets}
    return {**body, "digest": digest(canonical_json(body))}


def _contract_body(c: dict) -> dict:
    return {k: v for k, v in c.items() if k != "digest"}


def contract_digest_ok(c: dict) -> bool:
    return isinstance(c.get("digest"), str) and c["digest"] == digest(canonical_json(_contract_body(c)))


def _extends(recorded: dict, requested: dict) -> bool:
    """requested = recorded plus appended targets (same binary, reference and
    options, recorded targets as a prefix)."""
    a, b = _contract_body(recorded), _contract_body(requested)
    ta, tb = a.pop("targets"), b.pop("targets")
    return a == b and len(tb) >= len(ta) and tb[:len(ta)] == ta


@dataclass
class OutputSpec:
    key: str            # stable logical name, e.g. "frame/frame_000001.bin"
    path: Path
    rows: int = 0
    record_bytes: int = 0


# ---------------------------------------------------------------------------
# session
# ---------------------------------------------------------------------------
class OutputSession:
    """One multi-frame job in root.  outputs: the payloads the contract
    promises (in target order); metadata: the manifests, rewritten by the run
    and fingerprinted by the completion record.  Use:

        s = OutputSession(root, contract, outputs, metadata)
        s.begin()
        if not s.complete:
            for each frame: s.publish(spec, payload_bytes)   # in order, from frame 0
            s.publish_metadata(manifest_path, bytes)
            s.finish()
        s.close()
    """

    def __init__(self, root, contract: dict, outputs: list[OutputSpec], metadata: list, durability: str = "rename",
                 repair: bool = False, allowed: tuple = ()):
        if durability not in ("rename", "fsync"):
            raise ValueError("durability must be 'rename' or 'fsync'")
        self.root = Path(root)
        self.contract = contract
        self.outputs = list(outputs)
        self.metadata = [Path(m) for m in metadata]
        self.durability = durability
        self.repair = repair
        self.allowed = {str(a) for a in allowed}
        self.state = self.root / STATE_DIR
        self.complete = False
        self.existing_files = 0
        self.verified_files = 0
        self.verified_bytes = 0
        self.repaired_files = 0
        self.quarantine_directory: Path | None = None
        self.archived_residue = 0
        self._lock_fd = None
        self._pending = False
        self._verified: dict[str, dict] = {}
        self._published_appended = False

    # -- locking -------------------------------------------------------------
    def _lock(self):
        self.state.mkdir(parents=True, exist_ok=True)
        fd = os.open(self.state / LOCK, os.O_RDWR | os.O_CREAT, 0o644)
        try:
            fcntl.flock(fd, fcntl.LOCK_EX | fcntl.LOCK_NB)
        except OSError as e:
            os.close(fd)
            if e.errno in (errno.EWOULDBLOCK, errno.EAGAIN):
                raise IntegrityError("run directory already locked") from None
            raise
        self._lock_fd = fd

    def close(self) -> None:
        if self._lock_fd is not None:
            fcntl.flock(self._lock_fd, fcntl.LOCK_UN)
            os.close(self._lock_fd)
            self._lock_fd = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- helpers -------------------------------------------------------------
    def _intent_path(self, spec: OutputSpec) -> Path:
        return self.state / INTENTS / (spec.key.replace("/", "__") + ".json")

    def _quarantine(self, path: Path) -> None:
        if self.quarantine_directory is None:
            stamp = time.strftime("%Y%m%dT%H%M%SZ", time.gmtime())
            self.quarantine_directory = self.state / QUARANTINE / f"{stamp}-{os.getpid()}"
            self.quarantine_directory.mkdir(parents=True, exist_ok=True)
        dest = self.quarantine_directory / path.name
        k = 1
        while dest.exists():
            dest = self.quarantine_directory / f"{path.name}.{k}"
            k += 1
        os.rename(path, dest)

    def _check_intent(self, spec: OutputSpec):
        """The intent record of spec: dict, None (absent) or "invalid"."""
        ip = self._intent_path(spec)
        if not ip.exists():
            return None
        try:
            rec = read_json(ip)
        except (ValueError, UnicodeDecodeError):
            return "invalid"
        if (rec.get("schema") != INTENT_SCHEMA or rec.get("key") != spec.key or rec.get("file") != spec.path.name
                or not isinstance(rec.get("bytes"), int) or not _HEX32.match(str(rec.get("digest", "")))):
            return "invalid"
        return rec

    # -- begin ---------------------------------------------------------------
    def begin(self) -> None:
        """Validate the directory against the contract (rejections leave it
        unchanged), then archive crash residue and record the contract."""
        self._lock()
        try:
            self._begin()
        except BaseException:
            self.close()
            raise

    def _begin(self) -> None:
        cpath = self.state / CONTRACT
        recorded = None
        if cpath.exists():
            try:
                recorded = read_json(cpath)
            except (ValueError, UnicodeDecodeError) as e:
                raise IntegrityError(f"contract.json does not parse: {e}") from None
            if not contract_digest_ok(recorded):
                raise IntegrityError("contract.json is not a contract of this tool (digest mismatch)")
            if recorded["digest"] != self.contract["digest"]:
                if _extends(recorded, self.contract):
                    self._pending = True
                else:
                    a, b = _contract_body(recorded), _contract_body(self.contract)
                    if a["targets"] != b["targets"] and {k: v for k, v in a.items() if k != "targets"} == \
                            {k: v for k, v in b.items() if k != "targets"}:
                        ta, tb = a["targets"], b["targets"]
                        if len(tb) < len(ta) and tb == ta[:len(tb)]:
                            raise IntegrityError(
                                "resume rejected: targets changed/reordered; only an identical prefix plus append "
                                f"is allowed: the recorded run lists {len(ta)} targets, this command only {len(tb)} "
                                "(a shorter list is not a resume)")
                        i = next(i for i in range(min(len(ta), len(tb))) if ta[i] != tb[i])
                        raise IntegrityError(
                            "resume rejected: targets changed/reordered; only an identical prefix plus append is "
                            f"allowed: target {i} of the recorded {len(ta)} differs (recorded -> requested): "
                            + describe_json_differences(ta[i], tb[i]))
                    raise IntegrityError(
                        "resume rejected: run configuration/reference/binary changed (recorded -> requested): "
                        + describe_json_differences({k: v for k, v in a.items() if k != "targets"},
                                                    {k: v for k, v in b.items() if k != "targets"})
                        + "; use a new output directory for a different run")
        else:
            # a directory without a contract must not hold outputs of another run
            for spec in self.outputs:
                if spec.path.exists() or spec.path.is_symlink():
                    raise IntegrityError(f"pre-ledger output {spec.path.name} without a run record; "
                                         "use a fresh output directory")

        # payloads: verify what exists
        damaged: list[tuple[OutputSpec, str]] = []
        for spec in self.outputs:
            rec = self._check_intent(spec)
            present = spec.path.exists() or spec.path.is_symlink()
            if present:
                self.existing_files += 1
                if spec.path.is_symlink() or not spec.path.is_file():
                    raise IntegrityError(f"output {spec.path.name} is not a regular file")
            if rec == "invalid":
                damaged.append((spec, "invalid intent"))
                continue
            if present and rec is None:
                damaged.append((spec, "payload without an intent"))
                continue
            if present:
                fp = fingerprint(spec.path)
                if fp["bytes"] != rec["bytes"] or fp["digest"] != rec["digest"]:
                    damaged.append((spec, "payload does not match its intent"))
                    continue
                self.verified_files += 1
                self.verified_bytes += fp["bytes"]
                self._verified[spec.key] = rec
        # entries the ledger does not name
        names = {s.path.name for s in self.outputs} | {m.name for m in self.metadata} | {STATE_DIR} | self.allowed
        dirs = {s.path.parent for s in self.outputs} | {self.root}
        residue = []
        for d in dirs:
            for e in sorted(d.iterdir()):
                if e.name in names:
                    continue
                if _RESIDUE.match(e.name) and e.is_file() and not e.is_symlink():
                    residue.append(e)
                    continue
                if e.name.startswith(TMP_PREFIX):
                    continue  # reserved name, not ours to move: reported, never rejected
                raise IntegrityError(f"untracked entry in output directory: {e.name}")
        # metadata: must parse when present
        for m in self.metadata:
            if m.exists():
                if m.is_symlink() or not m.is_file():
                    raise IntegrityError(f"metadata {m.name} is not a regular file")
                try:
                    read_json(m)
                except (ValueError, UnicodeDecodeError):
                    damaged.append((OutputSpec(key="meta/" + m.name, path=m), "manifest does not parse"))
                else:
                    self.verified_files += 1
        if damaged:
            if not self.repair:
                what = "; ".join(f"{s.path.name}: {why}" for s, why in damaged[:8])
                raise IntegrityError(f"damaged output ({len(damaged)} files): {what}; rerun with repair "
                                     "(--repair) to quarantine and regenerate them")
            for spec, _ in damaged:
                if spec.path.exists():
                    self._quarantine(spec.path)
                ip = self._intent_path(spec)
                if ip.exists():
                    self._quarantine(ip)
                self.repaired_files += 1
                self._verified.pop(spec.key, None)
        for e in residue:
            self._quarantine(e)
            self.archived_residue += 1
        # record the contract (pending while an append has not published)
        (self.state / INTENTS).mkdir(parents=True, exist_ok=True)
        if recorded is None:
            atomic_write(self.state / CONTRACT, canonical_json(self.contract), True, self.durability)
        elif self._pending:
            atomic_write(self.state / PENDING, canonical_json(self.contract), True, self.durability)
        elif (self.state / PENDING).exists():
            self._quarantine(self.state / PENDING)
        # complete: every payload verified, completion current, manifests match it
        self.complete = self._completion_current() and len(self._verified) == len(self.outputs)

    def _completion_current(self) -> bool:
        p = self.state / COMPLETE
        if not p.exists():
            return False
        try:
            rec = read_json(p)
        except (ValueError, UnicodeDecodeError):
            return False
        if rec.get("contract_digest") != self.contract["digest"]:
            return False
        man = rec.get("manifests", {})
        for m in self.metadata:
            if not m.exists() or fingerprint(m) != man.get(m.name):
                return False
        return True

    # -- publication ---------------------------------------------------------
    def publish(self, spec: OutputSpec, data: bytes) -> str:
        """The frame's payload, in target order: "kept" (an existing payload
        verified and equal), "published" (intent, then payload).  A replayed
        frame whose bytes differ from its recorded payload is an error."""
        if self._lock_fd is None:
            raise IntegrityError("session not begun")
        d = digest(data)
        rec = self._verified.get(spec.key)
        if rec is not None:
            if rec["digest"] != d or rec["bytes"] != len(data):
                raise IntegrityError(f"replayed {spec.path.name} differs from the recorded payload "
                                     "(nondeterministic output or changed environment)")
            return "kept"
        intent = {"schema": INTENT_SCHEMA, "key": spec.key, "file": spec.path.name, "rows": int(spec.rows),
                  "record_bytes": int(spec.record_bytes), "bytes": len(data), "digest": d}
        atomic_write(self._intent_path(spec), canonical_json(intent), True, self.durability)
        if spec.path.exists():
            # an intent without its payload was replayed; a payload appearing
            # now was not there at begin
            raise IntegrityError(f"{spec.path.name} appeared during the run")
        atomic_write(spec.path, data, replace=False, durability=self.durability)
        self._verified[spec.key] = intent
        if self._pending and not self._published_appended:
            # the first appended payload is on disk: the longer list is the contract
            atomic_write(self.state / CONTRACT, canonical_json(self.contract), True, self.durability)
            os.unlink(self.state / PENDING)
            self._pending = False
            self._published_appended = True
        return "published"

    def publish_metadata(self, path, data: bytes) -> None:
        atomic_write(Path(path), data, True, self.durability)

    def finish(self) -> None:
        """Publish the completion record once every payload and manifest is
        on disk."""
        missing = [s.path.name for s in self.outputs if s.key not in self._verified]
        if missing:
            raise IntegrityError(f"finish with {len(missing)} outputs unpublished: {missing[:4]}")
        if self._pending:  # an append of zero new payloads cannot happen: targets grew
            atomic_write(self.state / CONTRACT, canonical_json(self.contract), True, self.durability)
            os.unlink(self.state / PENDING)
            self._pending = False
        rec = {"schema": COMPLETE_SCHEMA, "contract_digest": self.contract["digest"], "durability": self.durability,
               "frames": len(self.outputs), "manifests": {m.name: fingerprint(m) for m in self.metadata}}
        if self.durability == "fsync":
            for m in self.metadata:
                _fsync_dir(m.parent)
        atomic_write(self.state / COMPLETE, canonical_json(rec), True, self.durability)
        self.complete = True


# ---------------------------------------------------------------------------
# audit (--verify-only)
# ---------------------------------------------------------------------------
def audit_run_directory(root, layout) -> dict:
    """Read-only verdict on one multi-frame output directory.  layout(contract,
    root) -> (outputs: [OutputSpec], metadata: [Path], inputs: [(path,
    fingerprint)]) derives the promised files from a recorded contract.
    Nothing is created, moved or rewritten."""
    root = Path(root)
    rep = {"schema": AUDIT_SCHEMA, "status": "error", "complete": False, "exit_code": 1, "output_dir": str(root),
           "frames_expected": 0, "frames_verified": 0, "bytes_verified": 0, "problems": [], "warnings": [],
           "missing": [], "mismatched": [], "no_intent": [], "invalid_intent": [], "untracked_outputs": [],
           "untracked": [], "residue": [], "completion": {"present": False, "state": "missing"},
           "manifests": [], "inputs": {"status": "match", "changed": [], "missing": []},
           "pending_append_state": "none"}

    def done(status):
        rep["status"] = status
        rep["complete"] = status == "complete"
        rep["exit_code"] = 0 if status == "complete" else 1
        for k in ("missing", "mismatched", "no_intent", "invalid_intent", "untracked_outputs", "untracked",
                  "residue", "appended_published"):
            rep.setdefault(k, [])
            rep[k + "_count"] = len(rep[k])
            rep[k] = rep[k][:LIST_LIMIT]
        return rep

    if not root.exists():
        return done("not_found")
    state = root / STATE_DIR
    if not root.is_dir() or not (state / CONTRACT).exists():
        return done("not_a_run")
    lock_fd = None
    if (state / LOCK).exists():
        lock_fd = os.open(state / LOCK, os.O_RDONLY)
        try:
            fcntl.flock(lock_fd, fcntl.LOCK_SH | fcntl.LOCK_NB)
        except OSError:
            os.close(lock_fd)
            return done("locked")
    try:
        try:
            contract = read_json(state / CONTRACT)
            if not contract_digest_ok(contract):
                raise ValueError("digest mismatch")
            outputs, metadata, inputs = layout(contract, root)
        except Exception as e:  # an unusable contract
            rep["problems"].append(f"contract: {e}")
            return done("damaged")
        rep["contract"] = {"digest": contract["digest"], "options": contract.get("options"),
                           "binary": contract.get("binary"), "reference": contract.get("reference")}
        rep["frames_expected"] = len(outputs)
        damaged = False
        probe = OutputSession(root, contract, outputs, metadata)
        for spec in outputs:
            rec = probe._check_intent(spec)
            present = spec.path.exists() or spec.path.is_symlink()
            if rec == "invalid":
                rep["invalid_intent"].append(spec.path.name)
                damaged = True
            elif not present:
                rep["missing"].append(spec.path.name)
            elif rec is None:
                rep["no_intent"].append(spec.path.name)
                damaged = True
            elif spec.path.is_symlink() or not spec.path.is_file():
                rep["mismatched"].append(spec.path.name)
                damaged = True
            else:
                fp = fingerprint(spec.path)
                if fp["bytes"] != rec["bytes"] or fp["digest"] != rec["digest"]:
                    rep["mismatched"].append(spec.path.name)
                    damaged = True
                else:
                    rep["frames_verified"] += 1
                    rep["bytes_verified"] += fp["bytes"]
        names = {s.path.name for s in outputs} | {m.name for m in metadata} | {STATE_DIR}
        exts = {s.path.suffix for s in outputs}
        # an append cut off after its first new payload: the pending contract
        # accounts for the appended outputs that match their own intents
        pending = state / PENDING
        rep["appended_published"] = []
        if pending.exists():
            try:
                pc = read_json(pending)
                rep["pending_append_state"] = "append" if contract_digest_ok(pc) and _extends(contract, pc) else \
                    ("promoted" if pc == contract else "unrelated")
            except (ValueError, UnicodeDecodeError):
                rep["pending_append_state"] = "unreadable"
            if rep["pending_append_state"] == "append":
                p_outputs, _, _ = layout(pc, root)
                for spec in p_outputs[len(outputs):]:
                    rec = probe._check_intent(spec)
                    if spec.path.is_file() and not spec.path.is_symlink() and isinstance(rec, dict):
                        fp = fingerprint(spec.path)
                        if fp["bytes"] == rec["bytes"] and fp["digest"] == rec["digest"]:
                            rep["appended_published"].append(spec.path.name)
                            names.add(spec.path.name)
            rep["warnings"].append(f"contract.pending.json: {rep['pending_append_state']}")
        for d in {s.path.parent for s in outputs} | {root}:
            for e in sorted(d.iterdir()):
                if e.name in names:
                    continue
                if e.name.startswith(TMP_PREFIX):
                    rep["residue"].append(e.name)
                elif e.suffix in exts:
                    rep["untracked_outputs"].append(e.name)
                    damaged = True
                else:
                    rep["untracked"].append(e.name)
        comp = state / COMPLETE
        comp_rec = None
        if comp.exists():
            rep["completion"]["present"] = True
            try:
                comp_rec = read_json(comp)
                ok = _HEX32.match(str(comp_rec.get("contract_digest", "")))
                rep["completion"]["state"] = ("current" if comp_rec["contract_digest"] == contract["digest"]
                                              else "other_contract") if ok else "invalid"
                rep["completion"]["durability"] = comp_rec.get("durability")
            except (ValueError, UnicodeDecodeError, KeyError):
                rep["completion"]["state"] = "invalid"
                damaged = True
        manifests_ok = True
        for m in metadata:
            entry = {"file": m.name, "state": "ok"}
            if not m.exists():
                entry["state"] = "missing"
                manifests_ok = False
            else:
                try:
                    read_json(m)
                except (ValueError, UnicodeDecodeError):
                    entry["state"] = "invalid"
                    damaged = True
                else:
                    if rep["completion"]["state"] == "current" and \
                            fingerprint(m) != comp_rec.get("manifests", {}).get(m.name):
                        entry["state"] = "changed"
                        damaged = True
            rep["manifests"].append(entry)
        for path, fp in inputs:
            p = Path(path)
            if not p.exists():
                rep["inputs"]["missing"].append(str(path))
            elif fingerprint(p) != fp:
                rep["inputs"]["changed"].append(str(path))
        rep["inputs"]["status"] = ("changed" if rep["inputs"]["changed"] else
                                   "missing" if rep["inputs"]["missing"] else "match")
        if rep["residue"]:
            rep["warnings"].append(f"{len(rep['residue'])} temporary files left by an interrupted run")
        if damaged:
            return done("damaged")
        if rep["missing"] or rep["completion"]["state"] != "current" or not manifests_ok or \
                rep["appended_published"]:
            return done("incomplete")
        if rep["inputs"]["changed"]:
            return done("inputs_changed")
        return done("complete")
    finally:
        if lock_fd is not None:
            fcntl.flock(lock_fd, fcntl.LOCK_UN)
            os.close(lock_fd)
