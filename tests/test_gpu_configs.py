"""Parity of the HIP path with the FP64 CPU oracle on the configured
workloads that the smaller parity cases leave out:

  - the 128x128x64 MFISM stack at its configured size (p-Si, Boltzmann
    carriers, the nonlinear outer loop, the gate ramp);
  - Fermi-Dirac carriers (the reference's default, FerroX.cpp:556;
    ChargeDensity.cpp:4-11,71-81) with the bias driven into the degenerate
    range of either carrier;
  - devices of the 48-device batching workload (lateral extents 32..128,
    FE 3..10 nm) through the batched path, one by one against the oracle;
  - an FE block narrower than the domain with P_BC_flag branches 2, 3 and 4
    on its faces (DerivativeAlgorithm.H:45-467), vector P.

Bar: FP64 within 1e-10 relative (max-norm ratio) of the oracle, both sides
converged to 1e-13, with equal outer-iteration counts; FP32 + refinement
within 1e-6.
"""
import math

import numpy as np
import pytest

import ferrox_amd as fx
from ferrox_amd import configs
from ferrox_amd.batch import Batch
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10
FIELDS_SC = ("Pz", "Phi", "electrons", "holes", "charge")


def pair(text, **kw):
    kw.setdefault("mg_tol_rel", 1e-13)
    inp = fx.from_text(text, **kw)
    return fx.FerroX(inp), Oracle(inp)


def compare(g, o, names, tol, label=""):
    worst = 0.0
    for name in names:
        a, b = g.get(name), o.get(name)
        den = np.abs(b).max()
        e = float(np.abs(a - b).max() / (den if den > 0 else 1.0))
        print(label, name, e, flush=True)
        assert e < tol, (label, name, e)
        worst = max(worst, e)
    return worst


# ------------------------------------------------------------ 128x128x64 MFISM
@pytest.mark.parametrize("ramp", ["configured", "compressed"])
def test_mfism_128_configured_size(ramp):
    """128x128x64 cells over 64x64x32 nm: 26 nm p-Si (N_A = 1e24 m^-3,
    Boltzmann), 1 nm SiO2, 5 nm FE.  'configured' runs the first 3 steps of
    the 2000-step 0 -> 1 V ramp; 'compressed' reaches 1 V in those 3 steps,
    so the outer loop works on a strongly nonlinear charge."""
    n = 3
    text = configs.mfism_128(nsteps=2000) if ramp == "configured" else \
        configs.mfism(nsteps=n, ramp_steps=n, v_end=1.0)
    print(f"mfism_128 {ramp}: 128x128x64, {n} steps", flush=True)
    g, o = pair(text)
    sg, so = g.step(n), o.step(n)
    print(sg, so, flush=True)
    assert sg["steps_done"] == so["steps_done"] == n
    assert sg["total_outer_iters"] == so["total_outer_iters"]
    assert abs(sg["Phi_Bc_hi"] - so["Phi_Bc_hi"]) < 1e-15
    compare(g, o, FIELDS_SC, TOL, "mfism_128")
    g.close(); o.close()


# --------------------------------------------------------------- Fermi-Dirac
Q, KB, T = 1.602e-19, 1.38e-23, 300.0
NC, NV, EG, CHI = 2.8e25, 1.83e25, 1.12, 4.05


def etas(phi):
    kT = KB * T / Q
    phi_ref = CHI + 0.5 * EG + 0.5 * kT * math.log(NC / NV)
    eta_n = (phi - phi_ref + CHI) / kT
    return eta_n, -eta_n - EG / kT


@pytest.mark.parametrize("v_end,degenerate", [(3.0, "electrons"), (-2.0, "holes")])
def test_fermi_dirac_mfism(v_end, degenerate):
    """use_Fermi_Dirac = 1 on the MFISM stack, gate ramped to v_end in 5 steps:
    +3 V drives the electrons of the inversion layer to eta_n ~ 6.6, -2 V
    the accumulated holes to eta_p ~ 5.4, where FD_half departs from the
    Boltzmann exponential by orders of magnitude."""
    text = configs.mfism(nx=16, ny=16, nz=64, lx=8e-9, ly=8e-9, nsteps=5, ramp_steps=5, v_end=v_end)
    text += "\nuse_Fermi_Dirac = 1\n"
    print(f"fermi-dirac 16x16x64 v_end={v_end}", flush=True)
    g, o = pair(text)
    sg, so = g.step(5), o.step(5)
    print(sg, so, flush=True)
    assert sg["total_outer_iters"] == so["total_outer_iters"]
    compare(g, o, FIELDS_SC, TOL, "fd")
    sc = o.get("mask") >= 2
    en, ep = etas(o.get("Phi")[sc])
    eta = en if degenerate == "electrons" else ep
    assert eta.max() > 3.0, eta.max()
    # the carrier is far from its Boltzmann value there
    dens = o.get(degenerate)[sc] / (NC if degenerate == "electrons" else NV)
    i = int(np.argmax(eta))
    assert dens[i] < 0.2 * math.exp(eta[i])
    g.close(); o.close()


def test_fermi_dirac_mfism_streamed_fp32_refine():
    """Fermi-Dirac carriers under the precision form (FP32 V-cycles inside
    FP64 refinement; rho(phi) stays FP64): within 1e-6 of the oracle."""
    text = configs.mfism(nx=64, ny=16, nz=64, lx=32e-9, ly=8e-9, nsteps=3, ramp_steps=3, v_end=2.0)
    text += "\nuse_Fermi_Dirac = 1\n"
    g, o = pair(text, precision=1, mg_tol_rel=1e-10)
    g.step(3); o.step(3)
    compare(g, o, FIELDS_SC, 1e-6, "fd fp32")
    g.close(); o.close()


# ------------------------------------------------------------ batching form
def batch_extremes():
    """Devices of the 48-device batching workload (configs.batch_devices)
    covering its extremes: the smallest and the largest lateral extent and
    the thinnest and the thickest FE layer."""
    texts = configs.batch_devices(n=48, nsteps=10)
    infos = []
    for t in texts:
        inp = fx.from_text(t)
        nx, ny, nz = inp.n_cell
        infos.append((nx * ny, nz))
    pick = {int(np.argmin([a for a, _ in infos])), int(np.argmax([a for a, _ in infos])),
            int(np.argmin([z for _, z in infos])), int(np.argmax([z for _, z in infos]))}
    return [texts[i] for i in sorted(pick)]


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_devices_against_oracle(precision):
    texts = batch_extremes()
    dims = [tuple(fx.from_text(t).n_cell) for t in texts]
    print("batch devices", dims, "precision", precision, flush=True)
    nz = [d[2] for d in dims]
    assert min(nz) <= 16 + 13 and max(nz) >= 16 + 38          # FE ~3 nm and ~10 nm
    assert min(d[0] * d[1] for d in dims) <= 32 * 48 and max(d[0] * d[1] for d in dims) >= 128 * 96
    b = Batch([fx.from_text(t, precision=precision) for t in texts])
    b.step(10)
    tol = TOL if precision == 0 else 1e-6
    for dev, t in enumerate(texts):
        o = Oracle(fx.from_text(t, mg_tol_rel=1e-13))
        o.step(10)
        for name in ("Pz", "Phi"):
            a, r = b.get(dev, name), o.get(name)
            e = float(np.abs(a - r).max() / np.abs(r).max())
            print(dev, dims[dev], name, e, flush=True)
            assert e < tol, (dev, name, e)
        o.close()
    b.close()


# -------------------------------------------------- FE block, P_BC_flag 2/3/4
def fe_block_text(flags_lo, flags_hi):
    """32x24x36 cells, FE block narrower than the domain in x and y (and
    inside it in z) with the given P_BC flags on its faces; vector P with
    the coupling coefficients on; random initial P in the block."""
    return f"""
domain.prob_lo = -8.0e-9 -6.0e-9 0.0
domain.prob_hi = 8.0e-9 6.0e-9 9.0e-9
domain.n_cell = 32 24 36
domain.is_periodic = 1 1 0
boundary.lo = per per dir(0.0)
boundary.hi = per per dir(0.3)
prob_type = 2
TimeIntegratorOrder = 1
nsteps = 5
dt = 1.0e-13
Phi_Bc_lo = 0.0
Phi_Bc_hi = 0.3
random_seed = 9
Remnant_P = 2.0e-3 1.0e-3 3.0e-3
is_polarization_scalar = 0
SC_lo = -1.0 -1.0 -1.0
SC_hi = -1.0 -1.0 -1.0
DE_lo = -8.0e-9 -6.0e-9 0.0
DE_hi = 8.0e-9 6.0e-9 9.0e-9
FE_lo = -5.0e-9 -3.5e-9 3.0e-9
FE_hi = 4.0e-9 4.0e-9 7.5e-9
P_BC_flag_lo = {flags_lo}
P_BC_flag_hi = {flags_hi}
lambda = 3.0e-9
epsilon_0 = 8.85e-12
epsilonX_fe = 24.0
epsilonZ_fe = 24.0
epsilon_de = 10.0
epsilon_si = 11.7
alpha = -2.5e9
beta = 6.0e10
gamma = 1.5e11
alpha_12 = 1.0e8
alpha_112 = 2.0e9
alpha_123 = 3.0e9
BigGamma = 100
g11 = 1.0e-10
g44 = 1.0e-10
g12 = 2.0e-11
g44_p = 1.0e-11
"""


@pytest.mark.parametrize("flags_lo,flags_hi", [("2 4 0", "4 2 1"), ("3 2 4", "2 3 2"), ("4 3 2", "3 4 4")])
def test_fe_block_bc_flags(flags_lo, flags_hi):
    g, o = pair(fe_block_text(flags_lo, flags_hi))
    mask = o.get("mask")
    assert (mask[:, :, 0] != 0).all() and (mask[:, 0, :] != 0).all()      # FE narrower than the domain
    g.ComputePhi_Rho(); o.ComputePhi_Rho()
    g.CalculateTDGL_RHS(); o.CalculateTDGL_RHS()
    compare(g, o, ("PoissonRHS", "Phi", "GL_rhs_x", "GL_rhs_y", "GL_rhs_z"), 1e-11, f"flags {flags_lo}/{flags_hi}")
    g.step(5); o.step(5)
    compare(g, o, ("Px", "Py", "Pz", "Phi"), TOL, f"flags {flags_lo}/{flags_hi} 5 steps")
    g.close(); o.close()
