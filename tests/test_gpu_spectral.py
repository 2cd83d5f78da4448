"""The exact spectral solve that ends the V-cycle, against the V-cycle without it.

On a laterally periodic stack bounded in z by a Dirichlet face, the finest
coarse level whose planes fit one workgroup is solved exactly (plane
transforms on the matrix cores, one tridiagonal solve per lateral mode);
mg_exact_tail = 0 keeps the ordinary multigrid tail there.  Both V-cycles
converge to the same discrete solution, so at a relative tolerance of 1e-12
their phi must agree to 1e-9 of max|phi|, in FP64 and through the mixed
FP32 + refinement solve, and the exact solve must not need more V-cycles.

The cases cover lateral sizes that are multiples of the 16x16 matrix-core
tiles (the 64 x 64 stack) and sizes that are not (80 x 40: its exact level
is 40 x 20 x 9, so every transform runs with padded tiles and a partial
column block), a stack with a Neumann top face, and an exact level of 32
planes (the tridiagonal solves through global memory).
"""
import ctypes as C

import numpy as np
import pytest

import ferrox_amd as fx
from ferrox_amd import configs
from ferrox_amd._abi import FxSolveStats

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def neumann_top():
    text = configs.mfim(nx=48, ny=48, nz=36, lx=24e-9, ly=24e-9)
    return text.replace("boundary.hi = per per dir(0.0)", "boundary.hi = per per neu")


CASES = {
    "mfim_64": lambda: configs.mfim_64(nsteps=1),
    "padded_80x40": lambda: configs.mfim(nx=80, ny=40, nz=36, lx=40e-9, ly=20e-9),
    "neumann_top": neumann_top,
    # cubic cells: the exact level is 32 x 32 x 32, more planes than the
    # register-resident tridiagonal solve takes (16), so the general one runs
    "tall_32": lambda: configs.mfim(nx=64, ny=64, nz=64, lx=32e-9, ly=32e-9, lz=32e-9, fe_top=12e-9),
}


def solve(text, exact, precision, rhs):
    inp = fx.from_text(text, mg_tol_rel=1e-12, mg_max_iter=200, mg_exact_tail=exact, precision=precision,
                      poisson_direct=0)
    sim = fx.FerroX(inp)
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    phi = np.zeros(sim.shape)
    st = FxSolveStats()
    rc = sim._lib.fx_poisson_solve(sim._h, _dp(b), None, _dp(phi), C.byref(st))
    sim.close()
    assert rc == 0, rc
    return phi, st.vcycles


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", sorted(CASES))
def test_exact_solve_matches_multigrid_tail(case, precision):
    text = CASES[case]()
    sim = fx.FerroX(fx.from_text(text, poisson_direct=0))
    shape = sim.shape
    sim.close()
    rng = np.random.default_rng(7)
    rhs = rng.standard_normal(shape)
    phi1, n1 = solve(text, 1, precision, rhs)
    phi0, n0 = solve(text, 0, precision, rhs)
    scale = np.abs(phi0).max()
    assert scale > 0
    err = np.abs(phi1 - phi0).max() / scale
    print(f"{case} {precision}: vcycles exact {n1} tail {n0}, rel diff {err:.3e}")
    assert err <= 1e-9, err
    assert n1 <= n0, (n1, n0)
