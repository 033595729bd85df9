"""CPU restatement of the reference's frame policy (test side), used as the
oracle of the multi-frame driver:

  isBadPoiForFallback, shouldReplacePoiWithFallback,
  shouldTakeFrameFallbackResult, isPoiLevelFallbackCandidate   cli:493-558
  seedPoiQueueFromPrevious (temporal extrapolation off)         cli:560-624
  processFrame                                                  cli:3024-3210
  runMultiFrame baseline / threshold bookkeeping                cli:4247-4285
(cli = opencorr_cli_cuda_mvp.cpp).  FFTCC and IC-GN calls go to the C oracle.
"""
from __future__ import annotations

import math

import numpy as np

GOOD = np.float32(0.8)
MARGIN = np.float32(0.03)


def is_bad(q):
    z = q["result"][..., 2]
    return (z != -3.0) & (~np.isfinite(z) | (z < GOOD))


def should_replace(cur, fb):
    zf, zc = fb["result"][..., 2], cur["result"][..., 2]
    ok = np.isfinite(zf) & (zf != -3.0)
    return ok & (~np.isfinite(zc) | (zf >= zc))


def take_frame_fallback(skip, scratch):
    zs, zk = scratch["result"][..., 2], skip["result"][..., 2]
    minus3 = zk == -3.0
    a = np.isfinite(zs) & (zs != -3.0) & (zs >= GOOD)
    b = is_bad(skip) & should_replace(skip, scratch)
    return np.where(minus3, a, b)


def is_candidate(cur, prev, bad):
    z = cur["result"][..., 2]
    out = bad & (~np.isfinite(z) | (z < 0))
    rest = bad & ~out
    if prev is None:
        return out | rest
    same = (cur["x"] == prev["x"]) & (cur["y"] == prev["y"])
    zp = prev["result"][..., 2]
    rest_true = rest & (~same | ~np.isfinite(zp) | (zp >= GOOD))
    return out | rest_true


def seed_from_previous(template, prev):
    q = template.copy()
    zp = prev["result"][:, 2]
    ok = np.isfinite(zp) & (zp >= 0)
    q["deformation"][ok] = prev["deformation"][ok]
    q["result"][ok] = 0
    q["strain"][ok] = 0
    return q


class PolicyOracle:
    """The fixed-reference sequence with FFTCC cached on the template."""

    def __init__(self, oracle, ref, template, radius, conv=0.001, stop=10.0):
        self.o = oracle
        self.ref = ref.astype(np.float32)
        self.template = template
        self.r = radius
        self.conv, self.stop = conv, stop
        self.prepared = oracle.PreparedRef(self.ref, radius, template)
        self.prev = None
        self.baseline = None

    def _fftcc_icgn(self, tar, q, idx=None):
        self.prepared.compute(tar, q, idx)
        self.o.icgn(self.ref, tar, self.r, q, self.conv, self.stop)

    def frame(self, tar_img, skip_policy=True, poi_fallback=True):
        tar = tar_img.astype(np.float32)
        n = len(self.template)
        stats = {"skip_fftcc_used": 0, "fallback_used": 0, "poi_fallback_count": 0}
        use_skip = skip_policy and self.baseline is not None and self.prev is not None
        if not use_skip:
            q = self.template.copy()
            self._fftcc_icgn(tar, q)
            ratio = np.float32(is_bad(q).sum()) / np.float32(n) if n else np.float32(0)
            if skip_policy and self.baseline is None:
                self.baseline = ratio
        else:
            stats["skip_fftcc_used"] = 1
            q = seed_from_previous(self.template, self.prev)
            self.o.icgn(self.ref, tar, self.r, q, self.conv, self.stop)
            bad = is_bad(q)
            ratio = np.float32(bad.sum()) / np.float32(n)
            cand = np.nonzero(is_candidate(q, self.prev, bad))[0] if poi_fallback else np.zeros(0, int)
            if ratio > np.float32(self.baseline + MARGIN):
                stats["fallback_used"] = 1
                skip = q.copy()
                q = self.template.copy()
                self._fftcc_icgn(tar, q)
                take = take_frame_fallback(skip, q)
                q = np.where(take, q, skip)
            elif poi_fallback and len(cand):
                stats["poi_fallback_count"] = len(cand)
                fb = self.template[cand].copy()
                self._fftcc_icgn(tar, fb, cand)
                rep = should_replace(q[cand], fb)
                q[cand[rep]] = fb[rep]
        self.prev = q
        return q, stats
