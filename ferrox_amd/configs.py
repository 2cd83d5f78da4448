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
