"""Polycrystalline FE layer: the grain generator and the oracle's per-cell
rotation and Landau alpha (runs without a GPU).

The rotation check is independent of the reference formula's transcription:
R(alpha, beta, theta) of the use_Euler_angles = 1 convention
(TotalEnergyDensity.cpp:167-178) is the passive z-x-z product
Rz(theta) Rx(beta) Rz(alpha), composed here from elementary rotations.
"""
import numpy as np
import pytest
from scipy import stats

import ferrox_amd as fx
from ferrox_amd import configs
from oracle.oracle import Oracle


def rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])


def rx(b):
    c, s = np.cos(b), np.sin(b)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])


def R_of(a_deg, b_deg, t_deg):
    a, b, t = np.radians([a_deg, b_deg, t_deg])
    return rz(t) @ rx(b) @ rz(a)


def small_poly(nx=24, ny=20, nz=40, n_grains=9, seed=3, **kw):
    text = configs.mfim(nx=nx, ny=ny, nz=nz, lx=nx * 0.5e-9, ly=ny * 0.5e-9, lz=nz * 0.1e-9,
                        de_top=1.0e-9, fe_top=3.6e-9, nsteps=1, bias=0.5) + "\nuse_Euler_angles = 1\n"
    inp = fx.from_text(text, **kw)
    return inp, configs.polycrystal_fields(inp, n_grains=n_grains, seed=seed)


def test_rotations_are_proper_and_euler_round_trip():
    Q = configs.random_rotations(np.random.default_rng(0), 2000)
    eye = np.broadcast_to(np.eye(3), Q.shape)
    assert np.allclose(Q @ np.swapaxes(Q, 1, 2), eye, atol=1e-14)
    assert np.allclose(np.linalg.det(Q), 1.0, atol=1e-14)
    a, b, t = configs.euler_from_matrix(Q)
    for q in range(len(Q)):
        assert np.allclose(R_of(a[q], b[q], t[q]), Q[q], atol=1e-12)


def test_rotations_are_haar_uniform():
    """Under the Haar measure the polar axis (third row) is uniform on the
    sphere, so cos(beta) = R33 is U(-1, 1) and the azimuth is U(-pi, pi); the
    rotation angle has density (1 - cos w) / pi on [0, pi]."""
    Q = configs.random_rotations(np.random.default_rng(1), 20000)
    assert stats.kstest(Q[:, 2, 2], stats.uniform(loc=-1, scale=2).cdf).pvalue > 1e-3
    az = np.arctan2(Q[:, 2, 0], -Q[:, 2, 1])
    assert stats.kstest(az, stats.uniform(loc=-np.pi, scale=2 * np.pi).cdf).pvalue > 1e-3
    w = np.arccos(np.clip((np.trace(Q, axis1=1, axis2=2) - 1) / 2, -1, 1))
    cdf = lambda x: (x - np.sin(x)) / np.pi
    assert stats.kstest(w, cdf).pvalue > 1e-3


def test_voronoi_assignment_matches_brute_force():
    inp, f = small_poly()
    nx, ny, nz = inp.n_cell
    lo, hi = np.array(inp.prob_lo), np.array(inp.prob_hi)
    L = hi - lo
    h = L / np.array([nx, ny, nz])
    # regenerate the seed points the same way the generator draws them
    rng = np.random.default_rng(3)
    z0, z1 = inp.FE_lo[2], inp.FE_hi[2]
    pts = np.column_stack([rng.uniform(0, L[0], 9), rng.uniform(0, L[1], 9), rng.uniform(z0 - lo[2], z1 - lo[2], 9)])
    g = f["grain"]
    zc = lo[2] + (np.arange(nz) + 0.5) * h[2]
    fe_k = (zc >= inp.FE_lo[2]) & (zc <= inp.FE_hi[2])
    assert np.all(g[~fe_k] == -1) and np.all(g[fe_k] >= 0)
    for k in np.nonzero(fe_k)[0]:
        for j in range(ny):
            for i in range(nx):
                p = (np.array([i, j, k]) + 0.5) * h
                d = np.abs(pts - p)
                d[:, :2] = np.minimum(d[:, :2], L[:2] - d[:, :2])
                dist = np.sqrt((d * d).sum(1))
                best = np.argmin(dist)
                # ties aside (measure zero), the nearest seed point owns the cell
                assert g[k, j, i] == best or np.isclose(dist[g[k, j, i]], dist[best], rtol=1e-12)


def test_grain_phase_fraction_and_alpha():
    inp, f = small_poly(n_grains=40, nx=40, ny=40)
    g = f["grain"]
    la = f["landau_alpha"]
    per_grain = {int(q): np.unique(la[g == q]) for q in np.unique(g[g >= 0])}
    assert all(len(v) == 1 for v in per_grain.values())
    fe = sum(1 for v in per_grain.values() if v[0] == inp.alpha)
    pe = sum(1 for v in per_grain.values() if v[0] == 1.0e9)
    assert fe + pe == len(per_grain) == 40          # every grain owns cells here
    assert fe == 28                                  # round(0.7 * 40)
    assert np.all(la[g < 0] == inp.alpha)


def test_oracle_per_cell_landau_alpha():
    inp, f = small_poly(include_Grad=0, include_Elec=0)
    o = Oracle(inp)
    mask = o.get("mask")
    rng = np.random.default_rng(8)
    r = np.where(mask == 0, rng.uniform(-0.3, 0.3, mask.shape), 0.0)
    o.set("Pz", r)
    o.CalculateTDGL_RHS()
    base = o.get("GL_rhs_z")
    configs.apply_fields(o, f)
    assert np.array_equal(o.get("landau_alpha"), f["landau_alpha"])
    o.CalculateTDGL_RHS()
    gl = o.get("GL_rhs_z")
    al = f["landau_alpha"]
    expect = -inp.BigGamma * (al * r + inp.beta * r**3 + inp.gamma * r**5)
    sel = mask == 0
    assert np.allclose(gl[sel], expect[sel], rtol=1e-13, atol=1e-13 * np.abs(expect).max())
    fe = sel & (al == inp.alpha)
    assert np.array_equal(gl[fe], base[fe])
    assert not np.allclose(gl[sel & (al != inp.alpha)], base[sel & (al != inp.alpha)])


def test_oracle_rotated_gradient_on_quadratic_profile():
    """P_r a quadratic in position: every central difference is exact, so
    dF/dP_r = -g11 r3 H r3^T - g44 (r1 H r1^T + r2 H r2^T) with r_i the rows of
    the cell's rotation and H the constant Hessian."""
    inp, f = small_poly(include_Landau=0, include_Elec=0, g11=3.0e-10, g44=1.0e-10)
    o = Oracle(inp)
    configs.apply_fields(o, f)
    nx, ny, nz = inp.n_cell
    lo, hi = np.array(inp.prob_lo), np.array(inp.prob_hi)
    h = (hi - lo) / np.array([nx, ny, nz])
    z, y, x = np.meshgrid(*[lo[d] + (np.arange(n) + 0.5) * h[d] for d, n in ((2, nz), (1, ny), (0, nx))],
                          indexing="ij")
    c = np.array([1.0, -2.0, 3.0, 0.7, -1.3, 2.1]) * 1e15
    P = c[0] * x * x + c[1] * y * y + c[2] * z * z + c[3] * x * y + c[4] * y * z + c[5] * x * z
    H = np.array([[2 * c[0], c[3], c[5]], [c[3], 2 * c[1], c[4]], [c[5], c[4], 2 * c[2]]])
    mask = o.get("mask")
    o.set("Pz", np.where(mask == 0, P, 0.0))
    o.CalculateTDGL_RHS()
    gl = o.get("GL_rhs_z")
    g = f["grain"]
    zc = lo[2] + (np.arange(nz) + 0.5) * h[2]
    fe_k = np.nonzero((zc >= inp.FE_lo[2]) & (zc <= inp.FE_hi[2]))[0]
    checked = 0
    for k in fe_k[2:-2]:
        for j in range(1, ny - 1):
            for i in range(1, nx - 1):
                R = R_of(f["angle_alpha"][k, j, i], f["angle_beta"][k, j, i], f["angle_theta"][k, j, i])
                dF = -inp.g11 * R[2] @ H @ R[2] - inp.g44 * (R[0] @ H @ R[0] + R[1] @ H @ R[1])
                assert np.isclose(gl[k, j, i], -inp.BigGamma * dF, rtol=1e-7, atol=0), (k, j, i)
                checked += 1
    assert checked > 1000 and len(np.unique(g[g >= 0])) > 3
