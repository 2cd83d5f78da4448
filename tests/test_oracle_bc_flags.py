"""Pins of the oracle's masked stencils at every face of an FE block that is
narrower than the domain in x, y and z, for every P_BC_flag branch
(DerivativeAlgorithm.H:45-252 first derivatives, :257-467 second
derivatives).  Runs without a GPU.

Each branch stands for a boundary condition at the FE face x_f (s = x - x_f
along the direction under test); a profile that satisfies it is
differentiated exactly by that branch, so the check needs no copy of the
formula:

  flag 0  P = 0 on the face           s (a + b s)        dP = a + 2 b s,  d2P = 2 b
  flag 1  dP/dx = P / lambda          A (lambda + s)     dP = A,          d2P = 0
  flag 2  dP/dn = 0 on the face       c + s^2            dP = 2 s,        d2P = 2
  flag 3  stencil extends outside FE  a + b s + c s^2    dP = b + 2 c s,  d2P = 2 c
  flag 4  one-sided                   a + b s (first),   dP = b,          d2P = 2 c
                                      a + b s + c s^2 (second)

For flags other than 3 the cells outside the FE block hold large random
values: a branch that read across the face would fail.  First derivatives
are read from the Poisson RHS (-div P with one component set), second
derivatives from the gradient part of the TDGL right-hand side of Pz
(Landau and field terms off, g11 or g44 set to 1 to select the direction).
"""
import numpy as np
import pytest

import ferrox_amd as fx
from oracle.oracle import Oracle

H = (1.0e-9, 0.8e-9, 0.5e-9)
N = (16, 14, 18)
FE_CELLS = ((4, 11), (3, 10), (6, 13))     # inclusive cell ranges of the FE block
LAM = 3.0e-9
GAMMA = 100.0


def block_text(flags_lo, flags_hi, scalar):
    lo = [FE_CELLS[d][0] * H[d] for d in range(3)]
    hi = [(FE_CELLS[d][1] + 1) * H[d] for d in range(3)]
    L = [N[d] * H[d] for d in range(3)]
    return f"""
domain.prob_lo = 0.0 0.0 0.0
domain.prob_hi = {L[0]!r} {L[1]!r} {L[2]!r}
domain.n_cell = {N[0]} {N[1]} {N[2]}
domain.is_periodic = 1 1 0
boundary.lo = per per dir(0.0)
boundary.hi = per per dir(0.0)
prob_type = 2
TimeIntegratorOrder = 1
nsteps = 1
dt = 1.0e-13
Phi_Bc_lo = 0.0
Phi_Bc_hi = 0.0
random_seed = 3
Remnant_P = 0.0 0.0 1.0e-3
is_polarization_scalar = {1 if scalar else 0}
SC_lo = -1.0 -1.0 -1.0
SC_hi = -1.0 -1.0 -1.0
DE_lo = 0.0 0.0 0.0
DE_hi = {L[0]!r} {L[1]!r} {L[2]!r}
FE_lo = {lo[0]!r} {lo[1]!r} {lo[2]!r}
FE_hi = {hi[0]!r} {hi[1]!r} {hi[2]!r}
P_BC_flag_lo = {flags_lo[0]} {flags_lo[1]} {flags_lo[2]}
P_BC_flag_hi = {flags_hi[0]} {flags_hi[1]} {flags_hi[2]}
lambda = {LAM!r}
epsilon_0 = 8.85e-12
epsilonX_fe = 24.0
epsilonZ_fe = 24.0
epsilon_de = 10.0
epsilon_si = 11.7
alpha = -2.5e9
beta = 6.0e10
gamma = 1.5e11
alpha_12 = 0.0
alpha_112 = 0.0
alpha_123 = 0.0
BigGamma = {GAMMA!r}
g11 = 0.0
g44 = 0.0
g12 = 0.0
g44_p = 0.0
include_Landau = 0
include_Elec = 0
mg_tol_rel = 1e-12
"""


def coords(d):
    """Cell-centre coordinate along d, broadcast to (nz, ny, nx)."""
    x = (np.arange(N[d]) + 0.5) * H[d]
    shape = [1, 1, 1]
    shape[2 - d] = N[d]
    return np.broadcast_to(x.reshape(shape), (N[2], N[1], N[0]))


def fe_mask():
    m = np.zeros((N[2], N[1], N[0]), dtype=bool)
    (i0, i1), (j0, j1), (k0, k1) = FE_CELLS
    m[k0:k1 + 1, j0:j1 + 1, i0:i1 + 1] = True
    return m


def face_cells(d, side):
    """FE cells adjacent to the lower or upper FE face along d."""
    m = np.zeros((N[2], N[1], N[0]), dtype=bool)
    (i0, i1), (j0, j1), (k0, k1) = FE_CELLS
    sl = [slice(k0, k1 + 1), slice(j0, j1 + 1), slice(i0, i1 + 1)]
    c = FE_CELLS[d][0] if side == "lo" else FE_CELLS[d][1]
    sl[2 - d] = slice(c, c + 1)
    m[tuple(sl)] = True
    return m


def face_coord(d, side):
    return (FE_CELLS[d][0] if side == "lo" else FE_CELLS[d][1] + 1) * H[d]


# (profile, first derivative, second derivative) of s = x - x_f per flag;
# flag 4's first derivative is exact on linear profiles only
def profile(flag, s, order):
    a, b, c = 0.7, -1.3e9, 2.1e17
    if flag == 0:
        return s * (a * 1e9 + c * s), a * 1e9 + 2 * c * s, 2 * c + 0 * s
    if flag == 1:
        A = 5.0e6
        return A * (LAM + s), A + 0 * s, 0 * s
    if flag == 2:
        return a + c * s * s, 2 * c * s, 2 * c + 0 * s
    if flag == 3:
        return a + b * s + c * s * s, b + 2 * c * s, 2 * c + 0 * s
    if flag == 4:
        if order == 1:
            return a + b * s, b + 0 * s, 0 * s
        return a + b * s + c * s * s, b + 2 * c * s, 2 * c + 0 * s
    raise ValueError(flag)


def fill(flag, values):
    """Profile inside FE; outside, the profile itself for flag 3 (the branch
    reads it) and large random values otherwise (the branch must not)."""
    if flag == 3:
        return values
    rng = np.random.default_rng(11)
    return np.where(fe_mask(), values, rng.uniform(-1e3, 1e3, values.shape))


def flags_for(d, side, flag):
    # the face under test gets the flag; every other face flag 2, which
    # gives zero first and second derivatives of profiles constant across it
    lo, hi = [2, 2, 2], [2, 2, 2]
    (lo if side == "lo" else hi)[d] = flag
    return lo, hi


CASES = [(d, side, flag) for d in range(3) for side in ("lo", "hi") for flag in range(5)]


@pytest.mark.parametrize("d,side,flag", CASES)
def test_first_derivative_branches(d, side, flag):
    lo, hi = flags_for(d, side, flag)
    o = Oracle(fx.from_text(block_text(lo, hi, scalar=False)))
    s = coords(d) - face_coord(d, side)
    P, dP, _ = profile(flag, s, 1)
    comp = ("Px", "Py", "Pz")[d]
    for c in ("Px", "Py", "Pz"):
        o.set(c, fill(flag, P) if c == comp else np.zeros(P.shape))
    o.ComputePhi_Rho()
    rhs = o.get("PoissonRHS")
    cells = face_cells(d, side)
    # RHS = -div P in FE cells (R = I), 0 outside
    got, want = -rhs[cells], dP[cells]
    scale = np.abs(dP[fe_mask()]).max()
    assert np.abs(got - want).max() <= 1e-12 * scale, (got[:4], want[:4])
    assert np.all(rhs[~fe_mask()] == 0.0)


@pytest.mark.parametrize("d,side,flag", CASES)
def test_second_derivative_branches(d, side, flag):
    lo, hi = flags_for(d, side, flag)
    coef = {"g11": 1.0} if d == 2 else {"g44": 1.0}
    o = Oracle(fx.from_text(block_text(lo, hi, scalar=True), **coef))
    s = coords(d) - face_coord(d, side)
    P, _, d2P = profile(flag, s, 2)
    o.set("Pz", fill(flag, P))
    o.CalculateTDGL_RHS()
    gl = o.get("GL_rhs_z")
    cells = face_cells(d, side)
    # GL_z = Gamma (g11 d2P/dz2 + g44 (d2P/dx2 + d2P/dy2)) for scalar, unrotated P
    got, want = gl[cells] / GAMMA, d2P[cells]
    scale = max(np.abs(d2P).max(), abs(2.1e17))
    assert np.abs(got - want).max() <= 1e-9 * scale, (got[:4], want[:4])
    assert np.all(gl[~fe_mask()] == 0.0)
