"""Multi-frame driver: a fixed reference, one target per frame.

The reference, its gradients, the FFTCC reference spectra and the per-POI
Hessian records stay resident on the GPU; each frame uploads one target, and
the frame policy of opencorr_cli_cuda_mvp.cpp:3024-3210 (FFTCC + IC-GN on the
first frame; later frames seeded from the previous result, bad-ratio guard,
whole-frame fallback with per-POI merge, POI-level repair) runs on the device.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from ._native import POI_DTYPE, FrameStats
from .opencorr import Image2D, _check_queue, _Handle

SKIP_FFTCC = nat.FI_POLICY_SKIP_FFTCC
POI_FALLBACK = nat.FI_POLICY_POI_FALLBACK
LEARNED_SEED = nat.FI_POLICY_LEARNED_SEED
DEFAULT_POLICY = nat.FI_POLICY_DEFAULT


class FrameDriver(_Handle):
    _destroy = "fi_frame_destroy"

    def __init__(self, radius: int, conv_criterion: float = 0.001, stop_condition: float = 10,
                 device: int = 0):
        super().__init__()
        nat.call("fi_frame_create", int(radius), float(conv_criterion), float(stop_condition),
                 int(device), ctypes.byref(self._h))
        self.radius = int(radius)
        self.n = 0

    def set_reference(self, ref: Image2D) -> None:
        nat.call("fi_frame_set_ref", self._h, *ref.args())

    def set_pois(self, template: np.ndarray) -> None:
        q = _check_queue(template)
        nat.call("fi_frame_set_pois", self._h, nat.ptr(q), len(q))
        self.n = len(q)

    def prepare(self) -> dict:
        """Prepare the reference side now (if needed); the timing of the last
        preparation (fi_prepare_stats)."""
        st = nat.PrepareStats()
        nat.call("fi_frame_prepare", self._h, ctypes.byref(st))
        return st.as_dict()

    def upload_target(self, tar: Image2D) -> None:
        """Queue a target for the next run(); the copy to the device proceeds
        in the background (at most two targets queued)."""
        nat.call("fi_frame_upload_target", self._h, *tar.args())
        # the library reads the pixels until the frame that uses them starts
        self._queued = (getattr(self, "_queued", ()) + (tar,))[-3:]

    def run(self, policy: int = DEFAULT_POLICY) -> dict:
        st = FrameStats()
        nat.call("fi_frame_run", self._h, int(policy), ctypes.byref(st))
        return st.as_dict()

    def run_async(self, policy: int = DEFAULT_POLICY) -> None:
        """Launch a frame and return; wait() returns its statistics.  With
        graphs on, a frame that is not seeded from the previous one is queued
        on the device behind the frame still running (at most two in flight),
        so the device does not idle while the host waits, reads records and
        queues the next target."""
        nat.call("fi_frame_run_async", self._h, int(policy))

    def wait(self) -> dict:
        """Statistics of the oldest frame launched by run_async, once it has
        finished."""
        st = FrameStats()
        nat.call("fi_frame_wait", self._h, ctypes.byref(st))
        return st.as_dict()

    def download(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=POI_DTYPE)
        nat.call("fi_frame_download", self._h, nat.ptr(out), len(out))
        return out

    def records(self, age: int = 0, copy: bool = True) -> np.ndarray:
        """36-byte output records (x, y, u, v, ux, uy, vx, vy, zncc;
        rawio.RAW_DTYPE) of the last frame (age 0) or the one before (age 1,
        with set_records on), packed on the device.  copy=False returns a
        view of the driver's pinned buffer, valid until the second run after
        that frame."""
        from .rawio import RAW_DTYPE
        p = ctypes.c_void_p()
        m = ctypes.c_size_t(0)
        nat.call("fi_frame_records", self._h, int(age), ctypes.byref(p), ctypes.byref(m))
        if m.value == 0:
            return np.zeros(0, dtype=RAW_DTYPE)
        buf = (ctypes.c_char * (36 * m.value)).from_address(p.value)
        view = np.frombuffer(buf, dtype=RAW_DTYPE)
        return view.copy() if copy else view

    def set_records(self, on: bool = True) -> None:
        """Pack every frame's records on the device and copy them to host
        memory beside the next frame (records() then only waits)."""
        nat.call("fi_frame_set_records", self._h, int(bool(on)))

    def set_temporal_extrapolation(self, on: bool = True, max_delta: float = 0.0) -> None:
        """Two-frame extrapolation of the temporal seeds (--temporal-extrapolation,
        opencorr_cli_cuda_mvp.cpp:560-624); max_delta 0 = the driver's radius."""
        nat.call("fi_frame_set_temporal_extrapolation", self._h, int(bool(on)), float(max_delta))

    def download_redo(self) -> np.ndarray:
        """Per POI of the last frame: True where it went through FFTCC."""
        out = np.zeros(self.n, dtype=np.uint8)
        nat.call("fi_frame_download_redo", self._h, nat.ptr(out), len(out))
        return out.astype(bool)

    def set_shape(self, waves: int, rows_per_band: int) -> None:
        """IC-GN lane-to-POI shape: waves per POI and rows per band (0, 0 =
        default); it must fit the driver's subset."""
        nat.call("fi_frame_set_shape", self._h, int(waves), int(rows_per_band))

    def set_sampling(self, mode: str) -> None:
        """IC-GN target sampling: 'auto', 'table' or 'group' (see
        ICGN2D1.setSampling)."""
        modes = nat.SAMPLING_MODES
        if mode not in modes:
            raise ValueError(f"sampling must be one of {sorted(modes)}")
        nat.call("fi_frame_set_sampling", self._h, modes[mode])

    def set_seeder(self, seeder, zncc_threshold: float = 0.8) -> None:
        """Learned seeder for policy LEARNED_SEED (a seeder.LearnedSeeder, or
        None to detach).  POIs whose IC-GN result from the learned seed fails
        or has ZNCC below zncc_threshold are redone from FFTCC.  The seeder
        must stay open while it is attached."""
        self._seeder = seeder
        nat.call("fi_frame_set_seeder", self._h, seeder.handle if seeder is not None else None,
                 float(zncc_threshold))

    def set_seed_passes(self, passes: int) -> None:
        """Seeder passes per frame (1..4): each re-crops the target at the
        previous pass's rounded u, v."""
        nat.call("fi_frame_set_seed_passes", self._h, int(passes))

    def set_seed_consensus(self, on: bool = True, tau: float = 0.5, sample: str = "checker") -> None:
        """Neighbourhood consensus of the seeder's predictions after each pass
        (default on, tau 0.5 px, sample 'checker'): each seeded POI takes the mean of its seeded
        neighbours' predictions within +-radius, carried to it with their
        gradients, over the neighbours that agree within tau px.  sample: the
        POIs the network seeds ('all', 'checker' = a checkerboard of the
        consensus cells, 'quarter'); the others take their neighbours'
        consensus."""
        if sample not in nat.SEED_SAMPLES:
            raise ValueError(f"sample must be one of {sorted(nat.SEED_SAMPLES)}")
        nat.call("fi_frame_set_seed_consensus", self._h, int(bool(on)), float(tau), nat.SEED_SAMPLES[sample])

    def set_self_adaptive(self, on: bool = True) -> None:
        """Mixed subset radii: every template POI's (square) subset_radius is
        its FFTCC and IC-GN radius (DIC::setSelfAdaptive); call before
        set_pois takes effect on the next frame."""
        nat.call("fi_frame_set_self_adaptive", self._h, int(bool(on)))

    def set_graph(self, on: bool = True) -> None:
        """Replay each frame's device work as a captured hipGraph."""
        nat.call("fi_frame_set_graph", self._h, int(bool(on)))

    def set_persistent(self, blocks: int) -> None:
        """IC-GN persistent work-queue launch with `blocks` resident blocks
        (0 = one block per POI) for every radius class."""
        nat.call("fi_frame_set_persistent", self._h, int(blocks))

    def set_fused(self, on: bool = True) -> None:
        """One IC-GN launch over every radius class (per-POI radius) instead
        of one launch per class."""
        nat.call("fi_frame_set_fused", self._h, int(bool(on)))

    def autotune(self) -> list:
        """Time every IC-GN lane shape, sampling mode and launch mode per
        radius class on the uploaded target and keep the fastest; returns the
        timed configurations (dicts, 'chosen' marks the kept ones)."""
        cap = 512
        rep = (nat.TuneEntry * cap)()
        m = ctypes.c_size_t(0)
        nat.call("fi_frame_autotune", self._h, ctypes.cast(rep, ctypes.c_void_p), cap, ctypes.byref(m))
        return [rep[i].as_dict() for i in range(m.value)]

    def reset_sequence(self) -> None:
        nat.call("fi_frame_reset_sequence", self._h)

    def process(self, tar: Image2D, policy: int = DEFAULT_POLICY) -> tuple[np.ndarray, dict]:
        """Upload one target, run the frame, return (results, frame QA)."""
        self.upload_target(tar)
        stats = self.run(policy)
        return self.download(), stats
