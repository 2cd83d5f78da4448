"""Preconditioned conjugate gradients on semiconductor stacks (the direct
path's geometry with a Helmholtz term, fx_pcg.hip) against the FP64 CPU
oracle, whose multigrid solves the same systems.

  - MFISM steps (p-Si under SiO2 and FE, the nonlinear outer loop of
    ComputePhi_Rho) with Boltzmann and with Fermi-Dirac carriers: P, phi,
    n, p, rho within 1e-10 with equal outer-iteration counts, both sides
    converged to 1e-13;
  - fx_poisson_solve with a random, strongly varying Helmholtz term (up to
    a tenth of the Laplacian's diagonal): the plane-mean preconditioner is
    then far from the operator and CG still converges to the rule;
  - poisson_pcg = 0 keeps the multigrid, with the same answers;
  - the reference's stopping rule (1e-10) is met on the residual b - A phi.
"""
import numpy as np
import pytest

import ferrox_amd as fx
from ferrox_amd import configs
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def mfism_small(nsteps=5, fd=0, v_end=1.0):
    text = configs.mfism(nx=32, ny=32, nz=64, lx=16e-9, ly=16e-9, nsteps=nsteps, ramp_steps=nsteps, v_end=v_end)
    return text + f"\nuse_Fermi_Dirac = {fd}\n"


@pytest.mark.parametrize("fd", [0, 1])
def test_pcg_mfism_steps_match_oracle(fd):
    inp = fx.from_text(mfism_small(fd=fd), mg_tol_rel=1e-13)
    g, o = fx.FerroX(inp), Oracle(inp)
    sg, so = g.step(5), o.step(5)
    print("fd", fd, "pcg iterations", sg["total_vcycles"], "outer", sg["total_outer_iters"], so["total_outer_iters"],
          flush=True)
    assert sg["total_outer_iters"] == so["total_outer_iters"]
    assert 0 < sg["total_vcycles"] <= 30 * sg["total_outer_iters"]
    for name in ("Pz", "Phi", "electrons", "holes", "charge"):
        e = rel(g.get(name), o.get(name))
        print(name, e, flush=True)
        assert e < TOL, (name, e)
    g.close(); o.close()


def test_pcg_random_helmholtz_term():
    text = configs.mfim(nx=32, ny=64, nz=36, lx=16e-9, ly=32e-9, nsteps=1, bias=0.3)
    inp = fx.from_text(text, mg_tol_rel=1e-12)
    g, o = fx.FerroX(inp), Oracle(inp)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(g.shape) * 1e6
    a = -np.abs(rng.standard_normal(g.shape)) * 1e9 * 4
    pg, sg = g.poisson_solve(b, acoef=a)
    po, so = o.poisson_solve(b, acoef=a)
    print("pcg", sg, "oracle vcycles", so["vcycles"], flush=True)
    assert 0 < sg["vcycles"] < 60
    assert sg["mg_resid"] <= 1e-12 * max(sg["mg_bnorm"], sg["mg_resid0"])
    assert rel(pg, po) < TOL
    g.close(); o.close()


def test_pcg_off_keeps_the_multigrid():
    inp_p = fx.from_text(mfism_small(nsteps=3), mg_tol_rel=1e-13)
    inp_m = fx.from_text(mfism_small(nsteps=3), mg_tol_rel=1e-13, poisson_pcg=0)
    a, b = fx.FerroX(inp_p), fx.FerroX(inp_m)
    sa, sb = a.step(3), b.step(3)
    assert sa["total_outer_iters"] == sb["total_outer_iters"]
    assert sa["total_vcycles"] < sb["total_vcycles"]       # PCG iterations against V-cycles
    for name in ("Pz", "Phi", "electrons", "holes"):
        e = rel(a.get(name), b.get(name))
        print(name, "pcg vs multigrid", e, flush=True)
        assert e < TOL, (name, e)
    a.close(); b.close()


def test_pcg_reference_tolerance():
    """At the reference's 1e-10 the solve stops on the residual b - A phi, and
    the steps agree with the oracle's (whose multigrid stops elsewhere within
    the same rule) to the rule's accuracy."""
    inp = fx.from_text(mfism_small(nsteps=3), mg_tol_rel=1e-10)
    g, o = fx.FerroX(inp), Oracle(inp)
    sg, so = g.step(3), o.step(3)
    assert sg["total_outer_iters"] == so["total_outer_iters"]
    for name in ("Pz", "Phi", "electrons", "holes"):
        e = rel(g.get(name), o.get(name))
        print(name, e, flush=True)
        assert e < 1e-8, (name, e)
    g.close(); o.close()
