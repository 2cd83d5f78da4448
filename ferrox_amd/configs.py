"""Workload definitions as FerroX inputs text (reference key names).

mfim_64     64x64x36 cells over 32x32x9 nm: 4 nm Al2O3 (eps 10) under 5 nm
            HZO-like FE (eps 24), random Pz +-1e-3 C/m^2, zero bias, forward
            Euler dt = 1e-13 s, 1000 steps (the CPU-runnable case).
mfism_128   128x128x64 cells over 64x64x32 nm: 26 nm p-Si (N_A = 1e24 m^-3,
            Boltzmann carriers), 1 nm SiO2, 5 nm FE; gate ramped 0 -> 1 V over
            2000 steps.
mfim_256    256x256x128 cells, dx = dy = 0.5 nm, dz = 0.1 nm: 4 nm Al2O3,
            8 nm FE, 0.8 nm spacer (DE, the reference's default for cells in no
            region); 0.5 V bias, 500 steps.  The throughput workload.

Coefficients of the FE layer follow the shipped examples
(Exec/Examples/inputs_mfim_Noeb): alpha = -2.5e9, beta = 6e10, gamma = 1.5e11,
g11 = g44 = 1e-10, BigGamma = 100, P_BC_flag_lo = 3 3 0, P_BC_flag_hi = 3 3 1,
lambda = 3e-9.
"""

FE_COEFFS = """
alpha = -2.5e9
beta = 6.0e10
gamma = 1.5e11
alpha_12 = 0.0
alpha_112 = 0.0
alpha_123 = 0.0
BigGamma = 100
g11 = 1.0e-10
g44 = 1.0e-10
g12 = 0.0
g44_p = 0.0
P_BC_flag_lo = 3 3 0
P_BC_flag_hi = 3 3 1
lambda = 3.0e-9
epsilon_0 = 8.85e-12
epsilonX_fe = 24.0
epsilonZ_fe = 24.0
epsilon_si = 11.7
"""


def mfim(nx=64, ny=64, nz=36, lx=32e-9, ly=32e-9, lz=9e-9, de_top=4e-9, fe_top=9e-9,
         nsteps=1000, seed=1, bias=0.0) -> str:
    return f"""
domain.prob_lo = {-lx/2!r} {-ly/2!r} 0.0
domain.prob_hi = {lx/2!r} {ly/2!r} {lz!r}
domain.n_cell = {nx} {ny} {nz}
domain.max_grid_size = {nx} {ny} {nz}
domain.is_periodic = 1 1 0
boundary.lo = per per dir(0.0)
boundary.hi = per per dir({bias!r})
prob_type = 2
TimeIntegratorOrder = 1
nsteps = {nsteps}
dt = 1.0e-13
phi_tolerance = 0
Phi_Bc_lo = 0.0
Phi_Bc_hi = {bias!r}
random_seed = {seed}
Remnant_P = 0.0 0.0 1.0e-3
SC_lo = -1.0 -1.0 -1.0
SC_hi = -1.0 -1.0 -1.0
DE_lo = {-lx/2!r} {-ly/2!r} 0.0
DE_hi = {lx/2!r} {ly/2!r} {de_top!r}
FE_lo = {-lx/2!r} {-ly/2!r} {de_top!r}
FE_hi = {lx/2!r} {ly/2!r} {fe_top!r}
epsilon_de = 10.0
{FE_COEFFS}
"""


def mfim_64(nsteps=1000, seed=1) -> str:
    return mfim(nsteps=nsteps, seed=seed)


def mfism(nx=128, ny=128, nz=64, lx=64e-9, ly=64e-9, lz=32e-9, sc_top=26e-9, de_top=27e-9,
          nsteps=2000, ramp_steps=2000, v_end=1.0, seed=1) -> str:
    return f"""
domain.prob_lo = {-lx/2!r} {-ly/2!r} 0.0
domain.prob_hi = {lx/2!r} {ly/2!r} {lz!r}
domain.n_cell = {nx} {ny} {nz}
domain.is_periodic = 1 1 0
boundary.lo = per per dir(0.0)
boundary.hi = per per dir(0.0)
prob_type = 2
TimeIntegratorOrder = 1
nsteps = {nsteps}
dt = 1.0e-13
phi_tolerance = 0
Phi_Bc_lo = 0.0
Phi_Bc_hi = 0.0
bias_ramp_steps = {ramp_steps}
bias_ramp_start = 0.0
bias_ramp_end = {v_end!r}
random_seed = {seed}
Remnant_P = 0.0 0.0 1.0e-3
SC_lo = {-lx/2!r} {-ly/2!r} 0.0
SC_hi = {lx/2!r} {ly/2!r} {sc_top!r}
DE_lo = {-lx/2!r} {-ly/2!r} {sc_top!r}
DE_hi = {lx/2!r} {ly/2!r} {de_top!r}
FE_lo = {-lx/2!r} {-ly/2!r} {de_top!r}
FE_hi = {lx/2!r} {ly/2!r} {lz!r}
epsilon_de = 3.9
acceptor_doping = 1.0e24
donor_doping = 0.0
use_Fermi_Dirac = 0
{FE_COEFFS}
"""


def mfism_128(nsteps=2000) -> str:
    return mfism(nsteps=nsteps)


def mfim_256(nsteps=500, seed=1) -> str:
    return mfim(nx=256, ny=256, nz=128, lx=128e-9, ly=128e-9, lz=12.8e-9, de_top=4e-9, fe_top=12e-9,
                nsteps=nsteps, seed=seed, bias=0.5)


# --------------------------------------------------------------- polycrystal
def euler_from_matrix(Q):
    """Euler angles in degrees (angle_alpha, angle_beta, angle_theta) whose
    rotation matrix under use_Euler_angles = 1 (TotalEnergyDensity.cpp:167-178,
    R33 = cos(beta), R31 = sin(alpha) sin(beta), R32 = -cos(alpha) sin(beta),
    R13 = sin(beta) sin(theta), R23 = sin(beta) cos(theta)) equals Q.
    Q has shape (..., 3, 3) and is a proper rotation."""
    import numpy as np
    Q = np.asarray(Q, dtype=np.float64)
    b = np.arccos(np.clip(Q[..., 2, 2], -1.0, 1.0))
    a = np.arctan2(Q[..., 2, 0], -Q[..., 2, 1])
    t = np.arctan2(Q[..., 0, 2], Q[..., 1, 2])
    return np.degrees(a), np.degrees(b), np.degrees(t)


def random_rotations(rng, n):
    """n Haar-uniform rotations of SO(3) from uniformly random unit
    quaternions (normalised 4-D Gaussians), as (n, 3, 3) matrices."""
    import numpy as np
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)


def polycrystal_fields(inp, n_grains=400, fe_fraction=0.7, alpha_pe=1.0e9, seed=7):
    """Per-cell fields of a polycrystalline FE layer: the layer (cells whose
    centres lie inside FE_lo..FE_hi) is tiled by the Voronoi cells of
    n_grains seeded points, nearest point by physical distance with the
    periodic image in x and y.  Each grain takes a Haar-random rotation,
    written as the reference's Euler-angle fields in degrees
    (use_Euler_angles = 1), and is ferroelectric (Landau alpha from the inputs)
    or paraelectric (alpha_pe); round(fe_fraction * n_grains) grains are
    ferroelectric.  Cells outside the layer keep zero angles and the inputs'
    alpha.

    Returns {"angle_alpha", "angle_beta", "angle_theta", "landau_alpha",
    "grain"} as (nz, ny, nx) arrays (grain = -1 outside the layer)."""
    import numpy as np
    from scipy.spatial import cKDTree
    nx, ny, nz = inp.n_cell
    lo = np.array(inp.prob_lo, dtype=np.float64)
    hi = np.array(inp.prob_hi, dtype=np.float64)
    L = hi - lo
    h = L / np.array([nx, ny, nz])
    xc = lo[0] + (np.arange(nx) + 0.5) * h[0]
    yc = lo[1] + (np.arange(ny) + 0.5) * h[1]
    zc = lo[2] + (np.arange(nz) + 0.5) * h[2]
    inz = (zc >= inp.FE_lo[2]) & (zc <= inp.FE_hi[2])
    iny = (yc >= inp.FE_lo[1]) & (yc <= inp.FE_hi[1])
    inx = (xc >= inp.FE_lo[0]) & (xc <= inp.FE_hi[0])
    rng = np.random.default_rng(seed)
    z0, z1 = max(inp.FE_lo[2], lo[2]), min(inp.FE_hi[2], hi[2])
    pts = np.column_stack([rng.uniform(0.0, L[0], n_grains), rng.uniform(0.0, L[1], n_grains),
                           rng.uniform(z0 - lo[2], z1 - lo[2], n_grains)])
    Q = random_rotations(rng, n_grains)
    n_fe = int(round(fe_fraction * n_grains))
    is_fe = np.zeros(n_grains, dtype=bool)
    is_fe[rng.permutation(n_grains)[:n_fe]] = True
    ga, gb, gt = euler_from_matrix(Q)
    galpha = np.where(is_fe, inp.alpha, alpha_pe)

    # z is not periodic: a box far taller than the domain turns its wrap off
    tree = cKDTree(pts, boxsize=[L[0], L[1], 4.0 * L[2] + 1.0])
    grain = np.full((nz, ny, nx), -1, dtype=np.int32)
    X, Y = np.meshgrid(xc - lo[0], yc - lo[1], indexing="xy")
    sel2 = np.outer(iny, inx)
    for k in np.nonzero(inz)[0]:
        q = np.column_stack([X[sel2], Y[sel2], np.full(int(sel2.sum()), zc[k] - lo[2])])
        _, g = tree.query(q, k=1)
        layer = grain[k]
        layer[sel2] = g
    out = {"grain": grain}
    inside = grain >= 0
    gi = np.where(inside, grain, 0)
    out["angle_alpha"] = np.where(inside, ga[gi], 0.0)
    out["angle_beta"] = np.where(inside, gb[gi], 0.0)
    out["angle_theta"] = np.where(inside, gt[gi], 0.0)
    out["landau_alpha"] = np.where(inside, galpha[gi], inp.alpha)
    return out


def apply_fields(sim, fields):
    """Sets the per-cell fields of polycrystal_fields on a FerroX (or oracle)
    instance; P is left as initialised."""
    for key, field in (("angle_alpha", "alpha"), ("angle_beta", "beta"), ("angle_theta", "theta"),
                       ("landau_alpha", "landau_alpha")):
        sim.set(field, fields[key])


def mfim_poly_256(nsteps=500, seed=1) -> str:
    """The polycrystalline precision-form geometry: mfim_256 with the Euler
    convention of the grain rotations."""
    return mfim_256(nsteps=nsteps, seed=seed) + "\nuse_Euler_angles = 1\n"


# ------------------------------------------------------------------ batching
def batch_devices(n=48, seed=2025, nsteps=1000, sizes=(32, 48, 64, 96, 128)):
    """The batching workload: n independent MFIM devices, lateral extents
    drawn from `sizes` (cells of 0.5 nm, x and y independently), a 4 nm Al2O3
    layer (16 cells of 0.25 nm) under an FE layer of U(3, 10) nm (12-40
    cells), top bias U(-1, 1) V, per-device seed.  Returns inputs texts."""
    import numpy as np
    rng = np.random.default_rng(seed)
    texts = []
    for d in range(n):
        nx, ny = (int(v) for v in rng.choice(sizes, 2))
        nfe = int(round(rng.uniform(3e-9, 10e-9) / 0.25e-9))
        nz = 16 + nfe
        bias = float(rng.uniform(-1.0, 1.0))
        texts.append(mfim(nx=nx, ny=ny, nz=nz, lx=nx * 0.5e-9, ly=ny * 0.5e-9, lz=nz * 0.25e-9, de_top=4e-9,
                          fe_top=nz * 0.25e-9, nsteps=nsteps, seed=int(rng.integers(1, 2**31 - 1)), bias=bias))
    return texts
