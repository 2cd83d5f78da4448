"""The reference's own regression inputs through the drop-in path.

tests/golden/inputs/ holds the three inputs files of the reference's
regression harness (Exec/regression_inputs/{MIS/input_mis,
2D_MFIM/inputs_mfim_Noeb, 3D_MFIM/inputs_mfim_Noeb}), unchanged.  Each is
read by the library's ParmParse-compatible reader (fx_inputs_parse_file)
and run for its own nsteps on the GPU and on the FP64 CPU oracle:

  - MIS: 256x256x128 semiconductor + Channel (mask code 3), Fermi-Dirac
    carriers (the default), voltage sweep with Phi_Bc_inc = -0.05 and
    phi_tolerance = 5e-5 (a steady-state event raises the bias), no FE;
  - 2D_MFIM: 64x64x18 MFIM with prob_type = 1;
  - 3D_MFIM: the same stack with prob_type = 2.

Both sides solve to 1e-13 (the files' plot_int is kept: the plotfiles go to
a temporary directory), and P, phi, n, p, rho must agree to 1e-10 with
equal steps, sweep events and bias.  The two MFIM files also run at the
reference's own tolerance (1e-10), where both sides stop on the same V-cycles.
"""
import os

import numpy as np
import pytest

import ferrox_amd as fx
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "inputs")
TOL = 1e-10


def load(name, **over):
    inp = fx.parse_file(os.path.join(INPUTS, name))
    if over:
        fx.apply(inp, "\n".join(f"{k} = {v}" for k, v in over.items()))
    return fx.finalize(inp)


def rel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def run_pair(name, tmp_path, **over):
    inp = load(name, **over)
    print(f"{name}: n_cell {tuple(inp.n_cell)} nsteps {inp.nsteps} prob_type {inp.prob_type} "
          f"FD {inp.use_Fermi_Dirac} sweep {inp.voltage_sweep} tol {inp.mg_tol_rel}", flush=True)
    g = fx.FerroX(inp)
    g.set_plot_dir(str(tmp_path))
    o = Oracle(inp)
    sg, so = g.step(0), o.step(0)
    print("gpu", sg, flush=True)
    print("oracle", so, flush=True)
    return inp, g, o, sg, so


@pytest.mark.parametrize("name", ["2D_MFIM_inputs_mfim_Noeb", "3D_MFIM_inputs_mfim_Noeb"])
def test_mfim_regression_inputs(name, tmp_path):
    inp, g, o, sg, so = run_pair(name, tmp_path, mg_tol_rel=1e-13)
    assert inp.nsteps == 3 and sg["steps_done"] == so["steps_done"] == 3
    for f in ("Px", "Py", "Pz", "Phi", "Ez"):
        e = rel(g.get(f), o.get(f))
        print(name, f, e, flush=True)
        assert e < TOL, (f, e)
    # plot_int = 3: plotfiles at step 0 and step 3 (main.cpp:228-234, 563-569)
    assert sorted(os.listdir(tmp_path)) == ["plt00000000", "plt00000003"]
    g.close(); o.close()


@pytest.mark.parametrize("name", ["2D_MFIM_inputs_mfim_Noeb", "3D_MFIM_inputs_mfim_Noeb"])
def test_mfim_regression_inputs_reference_tolerance(name, tmp_path):
    """The multigrid (not the direct solve) at the file's own tolerance."""
    inp, g, o, sg, so = run_pair(name, tmp_path, poisson_direct=0)
    assert inp.mg_tol_rel == 1e-10
    assert sg["total_vcycles"] == so["total_vcycles"]
    for f in ("Pz", "Phi"):
        e = rel(g.get(f), o.get(f))
        print(name, f, e, flush=True)
        assert e < TOL, (f, e)
    g.close(); o.close()


def test_prob_type_1_initial_polarization():
    """prob_type = 1 (Initialization.cpp:88-105): the initial P of the 2D_MFIM
    file differs from prob_type = 2's and matches the oracle bit for bit."""
    a = load("2D_MFIM_inputs_mfim_Noeb")
    b = load("3D_MFIM_inputs_mfim_Noeb")
    ga, oa, gb = fx.FerroX(a), Oracle(a), fx.FerroX(b)
    for f in ("Px", "Py", "Pz"):
        assert np.array_equal(ga.get(f), oa.get(f)), f
    assert not np.array_equal(ga.get("Pz"), gb.get("Pz"))
    ga.close(); oa.close(); gb.close()


def test_mis_regression_input(tmp_path):
    inp, g, o, sg, so = run_pair("MIS_input_mis", tmp_path, mg_tol_rel=1e-13)
    assert inp.use_Fermi_Dirac == 1 and inp.voltage_sweep == 1
    mask = o.get("mask")
    assert (mask == 3).any()                                      # Channel cells
    for k in ("steps_done", "num_Vapp", "total_outer_iters", "stopped_steady"):
        assert sg[k] == so[k], k
    assert abs(sg["Phi_Bc_hi"] - so["Phi_Bc_hi"]) < 1e-15
    for f in ("Phi", "electrons", "holes", "charge", "Ez"):
        e = rel(g.get(f), o.get(f))
        print("MIS", f, e, flush=True)
        assert e < TOL, (f, e)
    g.close(); o.close()
