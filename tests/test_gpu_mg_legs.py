"""Fused multigrid kernels against the per-level kernels, V-cycle by V-cycle.

With mg_block_levels = 1 every small layered level below the finest runs its
down leg (sweeps, residual, restriction) and its up leg (prolongation,
sweeps) as one launch each, column blocks held in LDS with redundant halos;
with 0 the same level runs separate colour, restriction and prolongation
launches.  Both must give the same phi after each of the first V-cycles, to
rounding (the interior update multiplies by a per-plane 1/D where the
per-colour kernel divides): bar 1e-12 of max|phi|.  The cases cover the
coarsening patterns (1,1,1) and (1,1,0), periodic and non-periodic lateral
faces, and a Helmholtz term.

The same holds for the streaming sweeps of the large levels: two sweeps in
one pass (mg_fused_sweep = 2) against one sweep per pass (1) and against the
per-colour kernels (0), with the streaming kernels forced onto every level
they can serve (mg_stream_min_cells = 0), including tiles cut by the domain.
"""
import ctypes as C

import numpy as np
import pytest

import ferrox_amd as fx
from ferrox_amd import configs
from ferrox_amd._abi import FxSolveStats

pytestmark = pytest.mark.gpu

FX_ENOCONV_MG = 2


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def cycles(text, ncyc, block, rhs, acoef=None, **kw):
    """phi after exactly ncyc V-cycles from phi = 0 (a zero tolerance stops
    the solve at mg_max_iter)."""
    inp = fx.from_text(text, mg_tol_rel=0.0, mg_max_iter=ncyc, mg_block_levels=block, poisson_direct=0, **kw)
    sim = fx.FerroX(inp)
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    a = None if acoef is None else np.ascontiguousarray(acoef, dtype=np.float64)
    phi = np.zeros(sim.shape)
    st = FxSolveStats()
    rc = sim._lib.fx_poisson_solve(sim._h, _dp(b), None if a is None else _dp(a), _dp(phi), C.byref(st))
    assert rc == FX_ENOCONV_MG and st.vcycles == ncyc, (rc, st.vcycles)
    levels = sim.levels()
    sim.close()
    return phi, levels


def nonperiodic_mfim():
    text = configs.mfim(nx=64, ny=48, nz=36, lx=32e-9, ly=24e-9)
    text = text.replace("domain.is_periodic = 1 1 0", "domain.is_periodic = 0 0 0")
    text = text.replace("boundary.lo = per per dir(0.0)", "boundary.lo = neu dir(0.0) dir(0.0)")
    return text.replace("boundary.hi = per per dir(0.0)", "boundary.hi = dir(0.0) neu dir(0.0)")


def thin_dz():
    # dz = 0.1 nm: two z-only coarsenings (64 and 32 planes) before the
    # first full one, so the block legs serve levels of 64 and 32 planes
    return configs.mfim(nx=64, ny=64, nz=128, lx=32e-9, ly=32e-9, lz=12.8e-9, de_top=4e-9, fe_top=12e-9)


def thin_dz_256():
    # the same stack at 256 x 256: with the streaming kernels kept to levels
    # of 4M cells, the 256 x 256 x 32 level runs block legs of 8 x 8 columns
    # (the largest LDS footprint the legs take)
    return configs.mfim(nx=256, ny=256, nz=128, lx=128e-9, ly=128e-9, lz=12.8e-9, de_top=4e-9, fe_top=12e-9)


CASES = {
    "periodic_111_110": (lambda: configs.mfim_64(nsteps=1), {}),
    "nonperiodic": (nonperiodic_mfim, {}),
    "thin_dz": (thin_dz, {}),
    "thin_dz_256": (thin_dz_256, {"mg_stream_min_cells": 1 << 22}),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("helmholtz", [False, True])
def test_block_legs_match_per_level_kernels(case, helmholtz):
    make, kw = CASES[case]
    text = make()
    probe = fx.FerroX(fx.from_text(text, poisson_direct=0))
    shape = probe.shape
    probe.close()
    rng = np.random.default_rng(17)
    rhs = rng.standard_normal(shape) * 1e6
    acoef = None
    if helmholtz:
        # negative, as the Newton term d(rho)/d(phi) of the semiconductor
        # cells is: -a adds to the diagonal
        acoef = np.zeros(shape)
        acoef[: shape[0] // 3] = -1e2 * (1.0 + rng.random((shape[0] // 3,) + shape[1:]))
        acoef *= 8.854e-12 / (0.5e-9) ** 2
    for ncyc in (1, 3):
        p1, lv = cycles(text, ncyc, 1, rhs, acoef, **kw)
        p0, _ = cycles(text, ncyc, 0, rhs, acoef, **kw)
        scale = np.abs(p0).max()
        err = np.abs(p1 - p0).max() / scale
        print(case, helmholtz, ncyc, lv, err)
        assert scale > 0 and err < 1e-12, (ncyc, err)


def partial_tiles():
    # 96 x 20 columns: the last tile of 64 x 8 is cut in x and in y
    return configs.mfim(nx=96, ny=20, nz=36, lx=48e-9, ly=10e-9)


STREAM_CASES = {
    "periodic": lambda: configs.mfim_64(nsteps=1),
    "nonperiodic": nonperiodic_mfim,
    "partial_tiles": partial_tiles,
}


@pytest.mark.parametrize("case", sorted(STREAM_CASES))
@pytest.mark.parametrize("helmholtz", [False, True])
def test_two_sweep_pass_matches_single_sweeps(case, helmholtz):
    text = STREAM_CASES[case]()
    probe = fx.FerroX(fx.from_text(text, poisson_direct=0))
    shape = probe.shape
    probe.close()
    rng = np.random.default_rng(23)
    rhs = rng.standard_normal(shape) * 1e6
    acoef = None
    if helmholtz:
        acoef = np.zeros(shape)
        acoef[: shape[0] // 3] = -1e2 * (1.0 + rng.random((shape[0] // 3,) + shape[1:]))
        acoef *= 8.854e-12 / (0.5e-9) ** 2
    for ncyc in (1, 3):
        ref, lv = cycles(text, ncyc, 1, rhs, acoef, mg_fused_sweep=0)
        scale = np.abs(ref).max()
        for fused in (1, 2):
            p, _ = cycles(text, ncyc, 1, rhs, acoef, mg_fused_sweep=fused, mg_stream_min_cells=0)
            err = np.abs(p - ref).max() / scale
            print(case, helmholtz, ncyc, fused, lv, err)
            assert scale > 0 and err < 1e-12, (ncyc, fused, err)
