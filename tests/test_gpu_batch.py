"""Batching form on the GPU: every device of a batch against its own
single-device run (FP64 multigrid converged to 1e-13), graph replay against
eager launches, autotuned shapes against the defaults, FP32 + refinement
against FP64."""
import numpy as np
import pytest

import ferrox_amd as fx
from ferrox_amd import configs
from ferrox_amd.batch import Batch

pytestmark = pytest.mark.gpu


def small_texts(nsteps=10):
    # the last device has extents above 64 (two groups of m blocks per
    # transform tile on the matrix-core path)
    dims = [(32, 32, 28, 0.3), (48, 32, 30, -0.7), (32, 64, 36, 1.0), (64, 48, 44, -0.2), (128, 96, 32, 0.5)]
    return [configs.mfim(nx=nx, ny=ny, nz=nz, lx=nx * 0.5e-9, ly=ny * 0.5e-9, lz=nz * 0.25e-9, de_top=4e-9,
                         fe_top=nz * 0.25e-9, nsteps=nsteps, seed=7 + q, bias=v)
            for q, (nx, ny, nz, v) in enumerate(dims)]


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_batch_matches_single_device_runs():
    texts = small_texts()
    b = Batch([fx.from_text(t) for t in texts])
    b.step(10, graph=False)
    for d, t in enumerate(texts):
        g = fx.FerroX(fx.from_text(t, mg_tol_rel=1e-13))
        g.step(10)
        eP, ePhi = rel(b.get(d, "Pz"), g.get("Pz")), rel(b.get(d, "Phi"), g.get("Phi"))
        print(d, eP, ePhi)
        assert eP < 1e-10 and ePhi < 1e-10
        g.close()


def test_graph_replay_equals_eager_launches():
    texts = small_texts()
    a = Batch([fx.from_text(t) for t in texts])
    b = Batch([fx.from_text(t) for t in texts])
    a.step(7, graph=False)
    b.step(7, graph=True)
    for d in range(len(texts)):
        assert np.array_equal(a.get(d, "Pz"), b.get(d, "Pz"))
        assert np.array_equal(a.get(d, "Phi"), b.get(d, "Phi"))


def test_autotuned_shapes_give_the_same_result():
    texts = small_texts()
    a = Batch([fx.from_text(t) for t in texts])
    b = Batch([fx.from_text(t) for t in texts])
    rep = b.autotune()
    print(rep["chosen"])
    assert len(rep["stencil"]) == 12 and len(rep["spectral"]) == 12
    a.step(6)
    b.step(6)
    for d in range(len(texts)):
        assert np.array_equal(a.get(d, "Pz"), b.get(d, "Pz"))


def test_fp32_refinement_within_1e6_of_fp64():
    texts = small_texts(nsteps=60)
    a = Batch([fx.from_text(t) for t in texts])
    b = Batch([fx.from_text(t, precision=1) for t in texts])
    a.step(60)
    b.step(60)
    for d in range(len(texts)):
        eP, ePhi = rel(b.get(d, "Pz"), a.get(d, "Pz")), rel(b.get(d, "Phi"), a.get(d, "Phi"))
        print(d, eP, ePhi)
        assert eP < 1e-6 and ePhi < 1e-6


def test_devices_outside_the_batched_form_are_rejected():
    t = small_texts()[0].replace("domain.is_periodic = 1 1 0", "domain.is_periodic = 0 1 0").replace(
        "boundary.lo = per per dir(0.0)", "boundary.lo = neu per dir(0.0)")
    t = t.replace("boundary.hi = per per", "boundary.hi = neu per")
    with pytest.raises(fx.FerroXError):
        Batch([fx.from_text(t)])
