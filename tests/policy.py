"""CPU restatement of the reference's frame policy (test side), used as the
oracle of the multi-frame driver:

  isBadPoiForFallback, isEligibleTemporalSeedHistory,
  shouldReplacePoiWithFallback, shouldTakeFrameFallbackResult,
  isPoiLevelFallbackCandidate                                    cli:493-558
  seedPoiQueueFromPrevious (temporal extrapolation included)     cli:560-624
  processFrame                                                   cli:3024-3210
  runMultiFrame baseline / threshold / history bookkeeping       cli:4247-4285, 4376-4377
(cli = opencorr_cli_cuda_mvp.cpp).  FFTCC and IC-GN calls go to the C oracle.
Arithmetic that decides a branch is float32, as in the reference.
"""
from __future__ import annotations

import numpy as np

GOOD = np.float32(0.8)           # kSkipFftccZnccThreshold, cli:479
MARGIN = np.float32(0.03)        # kSkipFftccBadRatioMargin, cli:480
MIN_STEP = np.float32(0.05)      # kTemporalSeedMinStepPx, cli:481


def is_bad(q):
    z = q["result"][..., 2]
    return (z != -3.0) & (~np.isfinite(z) | (z < GOOD))


def eligible_history(q):
    """isEligibleTemporalSeedHistory (cli:499-504)."""
    d, z = q["deformation"], q["result"][..., 2]
    return np.isfinite(d[..., 0]) & np.isfinite(d[..., 6]) & np.isfinite(z) & (z >= GOOD)


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


def seed_from_previous(template, prev, prev2=None, extrapolate=False, max_delta=np.inf):
    """seedPoiQueueFromPrevious: (seeded queue, number of extrapolated seeds)."""
    q = template.copy()
    n = min(len(q), len(prev))
    zp = prev["result"][:n, 2]
    ok = np.zeros(len(q), bool)
    ok[:n] = np.isfinite(zp) & (zp >= 0)
    q["deformation"][ok] = prev["deformation"][:n][ok[:n]]
    count = 0
    if extrapolate and prev2 is not None and len(prev2) == len(prev):
        p, p2, t = prev[:n], prev2[:n], template[:n]
        same = (p["x"] == p2["x"]) & (p["y"] == p2["y"]) & (p["x"] == t["x"]) & (p["y"] == t["y"])
        du = (p["deformation"][:, 0] - p2["deformation"][:, 0]).astype(np.float32)
        dv = (p["deformation"][:, 6] - p2["deformation"][:, 6]).astype(np.float32)
        with np.errstate(over="ignore", invalid="ignore"):
            pu = (p["deformation"][:, 0] + du).astype(np.float32)
            pv = (p["deformation"][:, 6] + dv).astype(np.float32)
            step2 = (du * du).astype(np.float32) + (dv * dv).astype(np.float32)
            md = np.float32(max_delta)
            take = (ok[:n] & same & eligible_history(p) & eligible_history(p2) & np.isfinite(pu)
                    & np.isfinite(pv) & (step2 >= MIN_STEP * MIN_STEP)
                    & (np.abs(du) <= md) & (np.abs(dv) <= md))
        q["deformation"][:n, 0] = np.where(take, pu, q["deformation"][:n, 0])
        q["deformation"][:n, 6] = np.where(take, pv, q["deformation"][:n, 6])
        count = int(take.sum())
    q["result"][ok] = 0
    q["strain"][ok] = 0
    return q, count


def bad_ratio(q) -> np.float32:
    n = len(q)
    return np.float32(np.float32(is_bad(q).sum()) / np.float32(n)) if n else np.float32(0)


class PolicyOracle:
    """The fixed-reference sequence with FFTCC cached on the template.

    frame() returns (results, stats); stats restates the FrameProcessStats the
    reference reports per frame: skip_fftcc_used, fallback_used,
    poi_fallback_count, poi_fallback_recovered, temporal_seed_count,
    pre_repair_bad_ratio (the fallback trigger) and bad_ratio (after merge or
    repair); on a fallback frame stats["scratch"] holds the from-scratch pass.
    """

    def __init__(self, oracle, ref, template, radius, conv=0.001, stop=10.0, temporal=False,
                 self_adaptive=False, depth12="native"):
        """self_adaptive: every template POI's (square) subset_radius is its
        FFTCC and IC-GN radius (one cached FFTCC layout per radius), as the
        device driver's radius classes; radius is then the driver's radius
        (the temporal max_delta)."""
        self.o = oracle
        self.ref = oracle.as_f32(ref, depth12)
        self.depth12 = depth12
        self.template = template
        self.r = radius
        self.conv, self.stop = conv, stop
        self.temporal = temporal
        self.self_adaptive = self_adaptive
        radii = (template["subset_radius"][:, 0].astype(int) if self_adaptive
                 else np.full(len(template), radius))
        # per radius: template indices, position of each template index in
        # its class layout, and the class's cached reference spectra
        self.classes = []
        self.pos = np.zeros(len(template), dtype=np.int64)
        for rad in sorted(set(radii.tolist())):
            idx = np.nonzero(radii == rad)[0]
            self.pos[idx] = np.arange(len(idx))
            self.classes.append((rad, idx, oracle.PreparedRef(self.ref, rad, template[idx])))
        self.prev = None
        self.prev2 = None
        self.baseline = None
        self._grad = oracle.gradient(self.ref)
        self._coef_key = None

    def _icgn(self, tar, q):
        """ICGN2D1 on q with the reference gradients computed once and the
        target's B-spline table kept while the target stays the same."""
        key = (tar.shape, hash(tar.tobytes()))
        if key != self._coef_key:
            self._coef, self._coef_key = self.o.bspline(tar), key
        self.o.icgn(self.ref, tar, self.r, q, self.conv, self.stop, self_adaptive=self.self_adaptive,
                    prepared=(*self._grad, self._coef))

    def _fftcc_icgn(self, tar, q, idx=None):
        """FFTCC on the cached spectra, then IC-GN, of q = the template POIs
        idx (None: all of them)."""
        tidx = np.arange(len(self.template)) if idx is None else np.asarray(idx)
        for rad, cidx, prep in self.classes:
            m = np.isin(tidx, cidx)
            if not m.any():
                continue
            sub = q[m].copy()
            prep.compute(tar, sub, self.pos[tidx[m]])
            q[m] = sub
        self._icgn(tar, q)

    def learned_frame(self, tar_img, seeds, ok, threshold=0.8):
        """The learned-seed frame (FI_POLICY_LEARNED_SEED) from given seeds:
        IC-GN from the seeds; POIs that could not be seeded (ok false) or
        whose result is not a finite zncc >= threshold are redone from the
        template by FFTCC + IC-GN.  An unseeded POI takes the redo as it is
        (its record is the FFTCC-seeded path's); the others are replaced under
        shouldReplacePoiWithFallback (better zncc wins, cli:505-517)."""
        tar = self.o.as_f32(tar_img, self.depth12)
        q = seeds.copy()
        self._icgn(tar, q)
        z = q["result"][:, 2]
        cand = np.nonzero(~ok | ~(np.isfinite(z) & (z >= np.float32(threshold))))[0]
        if len(cand):
            fb = self.template[cand].copy()
            self._fftcc_icgn(tar, fb, cand)
            rep = should_replace(q[cand], fb) | ~ok[cand]
            q[cand[rep]] = fb[rep]
        return q, {"seed_fallback_count": len(cand)}

    def scratch(self, tar_img):
        """The frame from the template, FFTCC then IC-GN (--no-skip-fftcc)."""
        q = self.template.copy()
        self._fftcc_icgn(self.o.as_f32(tar_img, self.depth12), q)
        return q

    def frame(self, tar_img, skip_policy=True, poi_fallback=True, seed=None):
        """seed (optional): the previous frame's results to seed from instead
        of this oracle's own (one-step parity against a device sequence)."""
        tar = self.o.as_f32(tar_img, self.depth12)
        n = len(self.template)
        prev = self.prev if seed is None else seed
        st = {"skip_fftcc_used": 0, "fallback_used": 0, "poi_fallback_count": 0,
              "poi_fallback_recovered": 0, "temporal_seed_count": 0}
        use_skip = skip_policy and self.baseline is not None and prev is not None
        if not use_skip:
            q = self.template.copy()
            self._fftcc_icgn(tar, q)
            st["pre_repair_bad_ratio"] = st["bad_ratio"] = bad_ratio(q)
            if skip_policy and self.baseline is None:
                self.baseline = st["bad_ratio"]
        else:
            st["skip_fftcc_used"] = 1
            q, st["temporal_seed_count"] = seed_from_previous(self.template, prev, self.prev2, self.temporal,
                                                              np.float32(self.r))
            self._icgn(tar, q)
            bad = is_bad(q)
            n_bad = int(bad.sum())
            st["pre_repair_bad_ratio"] = st["bad_ratio"] = bad_ratio(q)
            cand = np.nonzero(is_candidate(q, prev, bad))[0] if poi_fallback else np.zeros(0, int)
            if st["pre_repair_bad_ratio"] > np.float32(self.baseline + MARGIN):
                st["fallback_used"] = 1
                skip = q.copy()
                q = self.template.copy()
                self._fftcc_icgn(tar, q)
                st["scratch"] = q.copy()
                take = take_frame_fallback(skip, q)
                q = np.where(take, q, skip)
                st["bad_ratio"] = bad_ratio(q)
            elif poi_fallback and len(cand):
                st["poi_fallback_count"] = len(cand)
                fb = self.template[cand].copy()
                self._fftcc_icgn(tar, fb, cand)
                rep = should_replace(q[cand], fb)
                st["poi_fallback_recovered"] = int((rep & ~is_bad(fb)).sum())
                q[cand[rep]] = fb[rep]
                st["bad_ratio"] = np.float32(np.float32(n_bad - st["poi_fallback_recovered"]) / np.float32(n))
        st["temporal_seed_used"] = int(st["temporal_seed_count"] > 0)
        self.prev2 = self.prev if seed is None else seed
        self.prev = q
        return q, st
