"""The TDGL update fused with the next step's Poisson RHS (k_tdgl_rhs_stream)
against the separate TDGL and RHS passes and against the FP64 CPU oracle.

The fused kernel does the same arithmetic as k_tdgl<true> and k_rhs_newton
(the compiler may contract multiply-adds differently in the two kernels), so
P, phi and PoissonRHS must agree to rounding (1e-13) with tdgl_fused_rhs = 0
and 1 (interior cells read their stencils from LDS, cells next to a face use the
masked stencils; partial 30 x 14 tiles at the domain edges; FE planes only).
PoissonRHS after a step is the RHS of the last solve (the reference's
plotfile variable), not the one prepared for the next step.
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


def fe_block(nx=40, ny=36, nz=30):
    """FE block narrower than the domain (lateral FE faces, masked stencils
    with flags 2/4 laterally and 0/1 in z), periodic sides."""
    text = configs.mfim(nx=nx, ny=ny, nz=nz, lx=nx * 0.5e-9, ly=ny * 0.5e-9, lz=nz * 0.25e-9, de_top=2e-9,
                        fe_top=6e-9, nsteps=8, bias=0.4)
    # later definitions win (ParmParse)
    return text + "\nFE_lo = -6.0e-9 -5.0e-9 2.0e-9\nFE_hi = 5.0e-9 6.0e-9 6.0e-9\n" \
                  "P_BC_flag_lo = 2 4 0\nP_BC_flag_hi = 4 2 1\n"


CASES = {
    "mfim_64": (configs.mfim(nx=64, ny=64, nz=36, nsteps=8, bias=0.2), None),
    "poly": (configs.mfim(nx=64, ny=48, nz=48, lx=32e-9, ly=24e-9, lz=4.8e-9, de_top=1.2e-9, fe_top=4.0e-9,
                          nsteps=8, bias=0.5) + "\nuse_Euler_angles = 1\n", "poly"),
    "fe_block": (fe_block(), None),
}


def make(text, poly, **kw):
    inp = fx.from_text(text, mg_tol_rel=1e-13, **kw)
    g = fx.FerroX(inp)
    f = None
    if poly:
        f = configs.polycrystal_fields(inp, n_grains=600 if poly == "poly_600" else 12, seed=3)
        configs.apply_fields(g, f)
    return inp, g, f


@pytest.mark.parametrize("case", list(CASES))
def test_fused_equals_separate_passes(case):
    text, poly = CASES[case]
    _, a, _ = make(text, poly, tdgl_fused_rhs=1)
    _, b, _ = make(text, poly, tdgl_fused_rhs=0)
    a.step(8)
    b.step(8)
    for name in ("Pz", "Phi", "PoissonRHS"):
        e = rel(a.get(name), b.get(name))
        print(case, name, "fused vs separate", e, flush=True)
        assert e < 1e-13, (name, e)
    a.close(); b.close()


@pytest.mark.parametrize("case", list(CASES))
def test_fused_matches_oracle(case):
    text, poly = CASES[case]
    inp, g, f = make(text, poly)
    o = Oracle(inp)
    if f is not None:
        configs.apply_fields(o, f)
    g.step(8)
    o.step(8)
    for name in ("Pz", "Phi", "PoissonRHS", "Ez"):
        e = rel(g.get(name), o.get(name))
        print(case, name, e, flush=True)
        assert e < 1e-10, (name, e)
    g.close(); o.close()


def test_fused_then_host_writes():
    """Writes between steps (P from the host, a Poisson solve through the
    context's buffers) invalidate the prepared RHS; the run stays equal to
    the unfused one."""
    text, _ = CASES["mfim_64"]
    _, a, _ = make(text, None, tdgl_fused_rhs=1)
    _, b, _ = make(text, None, tdgl_fused_rhs=0)
    rng = np.random.default_rng(4)
    for sim in (a, b):
        sim.step(3)
    P = a.get("Pz")
    P2 = P + np.where(a.get("mask") == 0, 1e-4 * rng.standard_normal(P.shape), 0.0)
    bb = rng.standard_normal(P.shape) * 1e5
    for sim in (a, b):
        sim.set("Pz", P2)
        sim.step(2)
        sim.poisson_solve(bb)
        sim.step(3)
    for name in ("Pz", "Phi", "PoissonRHS"):
        e = rel(a.get(name), b.get(name))
        print("host writes", name, e, flush=True)
        assert e < 1e-13, (name, e)
    a.close(); b.close()


STREAM_CASES = dict(CASES, poly_600=(CASES["poly"][0], "poly_600"))


@pytest.mark.parametrize("listed", [1, 2, 3])
@pytest.mark.parametrize("case", list(STREAM_CASES))
def test_streamed_update_equals_gathered(case, listed):
    """k_tdgl_stream and k_rhs_stream (operands through LDS) against k_tdgl<true>
    and k_rhs_newton (gathered): uniform rotation (mfim_64, fe_block), a
    12-grain rotation table staged in LDS (poly) and a 600-grain one, longer
    than the LDS copy, read from global memory by the update (poly_600); the
    cells next to a face by the streamed kernels over the planes holding them
    (masked stencils on the LDS window; tdgl_stream = 1 when those planes are
    few, as the FE planes next to the dielectric, and 3 always: fe_block has
    lateral FE faces with P_BC_flag 2 and 4, whose two-cell stencils reach
    past the window) or by list (tdgl_stream = 2, and 1 for fe_block)."""
    text, poly = STREAM_CASES[case]
    _, a, _ = make(text, poly, tdgl_stream=listed)
    _, b, _ = make(text, poly, tdgl_stream=0)
    a.step(8)
    b.step(8)
    for name in ("Pz", "Phi", "PoissonRHS"):
        e = rel(a.get(name), b.get(name))
        print(case, name, "streamed vs gathered", e, flush=True)
        assert e < 1e-13, (name, e)
    a.close(); b.close()
