"""FE cells next to a z face only (material bits MAT_ZFACE_LO / MAT_ZFACE_HI):
the streamed TDGL update and Poisson RHS form their masked z stencils
(DerivativeAlgorithm.H: DoubleDPDz and the DPDz of the mixed entries) as
linear forms inside the interior passes, instead of the face-plane launch.

Geometry: an FE layer between a DE layer below and a DE spacer above
(Initialization.cpp:169-206: unassigned cells are DE), periodic sides, so
every FE cell of the first and last FE plane is such a cell.  For each pair
of z flags that takes this path (0..3 on either side) and for flag 4 (two-cell
stencils, which keeps the face launch), with a uniform rotation and a
12-grain rotation table: the streamed passes (tdgl_stream = 1) against the
gathered kernels, which evaluate the masked stencils themselves
(tdgl_stream = 0), to 1e-13 (the linear forms reassociate the reference's
expressions), and against the FP64 oracle to 1e-10 after 8 steps.
"""
import numpy as np
import pytest

import ferrox_amd as fx
from ferrox_amd import configs
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def sandwich(lo, hi, grains):
    text = configs.mfim(nx=32, ny=32, nz=40, lx=16e-9, ly=16e-9, lz=10e-9, de_top=3e-9, fe_top=8e-9, nsteps=8,
                        bias=0.3)
    text += f"\nP_BC_flag_lo = 3 3 {lo}\nP_BC_flag_hi = 3 3 {hi}\n"
    if grains:
        text += "use_Euler_angles = 1\n"
    return text


def make(text, grains, **kw):
    inp = fx.from_text(text, mg_tol_rel=1e-13, **kw)
    g = fx.FerroX(inp)
    f = None
    if grains:
        f = configs.polycrystal_fields(inp, n_grains=grains, seed=5)
        configs.apply_fields(g, f)
    return inp, g, f


FLAGS = [(0, 1), (1, 0), (2, 3), (3, 2), (1, 2), (4, 4)]


@pytest.mark.parametrize("grains", [0, 12])
@pytest.mark.parametrize("lo,hi", FLAGS)
def test_zface_streamed_equals_gathered_and_oracle(lo, hi, grains):
    text = sandwich(lo, hi, grains)
    print("flags", lo, hi, "grains", grains, flush=True)
    inp, a, f = make(text, grains, tdgl_stream=1)
    _, b, _ = make(text, grains, tdgl_stream=0)
    o = Oracle(inp)
    if f is not None:
        configs.apply_fields(o, f)
    a.step(8)
    b.step(8)
    o.step(8)
    for name in ("Pz", "Phi", "PoissonRHS"):
        e = rel(a.get(name), b.get(name))
        eo = rel(a.get(name), o.get(name))
        print(lo, hi, grains, name, "streamed vs gathered", e, "vs oracle", eo, flush=True)
        assert e < 1e-13, (name, e)
        assert eo < 1e-10, (name, eo)
    a.close(); b.close(); o.close()
