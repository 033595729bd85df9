"""Recycled device and pinned buffers (the block pool).

A closed driver or engine hands its buffers and streams to the next one
created on the device.  Those buffers hold the previous owner's data, so
every result must come only from what the new owner writes: a frame
sequence, its device-packed records and the FFTCC2D / ICGN2D1 results of a
driver created after another one (on a different scene, different sizes)
was closed are byte-identical to those of the first driver in the process.
"""
import numpy as np
import pytest

import fasticgn_amd as fa
from fasticgn_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def frames(sc, tars, pois=None):
    out = []
    with fa.FrameDriver(15) as d:
        d.set_records(True)
        if pois is not None:
            d.set_self_adaptive(True)
        d.set_reference(fa.Image2D(sc.ref))
        d.set_pois(sc.pois if pois is None else pois)
        for t in tars:
            res, st = d.process(fa.Image2D(t))
            out.append((res.tobytes(), d.records().tobytes(),
                        tuple(st[k] for k in ("skip_fftcc_used", "valid_count", "bad_ratio", "iteration_mean"))))
    return out


def engines(sc):
    q = sc.pois.copy()
    with fa.FFTCC2D(15, 15) as f:
        f.setImages(fa.Image2D(sc.ref), fa.Image2D(sc.tar))
        f.compute(q)
    with fa.ICGN2D1(15, 15, 0.001, 10) as ic:
        ic.setImages(fa.Image2D(sc.ref), fa.Image2D(sc.tar))
        ic.prepare()
        ic.compute(q)
    return q.tobytes()


def test_recycled_buffers_give_identical_results():
    sc = synth.config0()
    tars = [sc.tar, np.roll(sc.tar, 1, axis=1), sc.tar]
    other = synth.affine_scene("pool-dirty", 640, 480, 9000, 20261206, t=(3.3, -2.2), grad=(2e-3, 0, 0, -1e-3),
                               radius=12, grid=7)
    say("scene", sc.ref.shape, len(sc.pois), "POIs; dirtying scene", other.ref.shape, len(other.pois), "POIs")
    first = frames(sc, tars)
    first_e = engines(sc)
    # dirty the kept blocks: the same sizes with other data (those blocks go
    # to the next driver), then other sizes and mixed radii
    same = synth.config0(seed=77)
    assert same.ref.shape == sc.ref.shape and len(same.pois) == len(sc.pois)
    frames(same, [same.tar, np.roll(same.tar, -2, axis=0)])
    engines(same)
    mixed = other.pois.copy()
    mixed["subset_radius"][:, 0] = mixed["subset_radius"][:, 1] = np.array([8, 12, 16], np.float32)[
        np.arange(len(mixed)) % 3]
    frames(other, [other.tar], pois=mixed)
    engines(other)
    for rnd in range(2):
        again = frames(sc, tars)
        for k, (a, b) in enumerate(zip(first, again)):
            assert a[0] == b[0], f"round {rnd} frame {k}: POIs differ"
            assert a[1] == b[1], f"round {rnd} frame {k}: records differ"
            assert a[2] == b[2], f"round {rnd} frame {k}: stats differ"
        assert engines(sc) == first_e, f"round {rnd}: FFTCC2D / ICGN2D1 results differ"
    assert first[0][2][1] > 0.95 * len(sc.pois)
