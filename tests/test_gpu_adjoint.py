"""Differentiation form on the GPU: the discrete adjoint against central
finite differences of the GPU forward model and of the FP64 CPU oracle.

Bar: adjoint gradients within 1e-5 relative of central differences
(h = 1e-6 relative) for every parameter (alpha, beta, gamma, g11, g44,
epsilonX_fe) and for entries of the gate waveform; the curve of the GPU
forward model within 1e-10 of the oracle's; gradients independent of the
checkpoint interval.  Solves are converged to 1e-13 so that solver error does
not enter the differences.
"""
import numpy as np
import pytest

import ferrox_amd as fx
from ferrox_amd import configs
from ferrox_amd.diff import central_fd, pv_curve, triangular_sweep
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

NSTEPS = 40
FD_TOL = 1e-5


def small_text(nsteps=NSTEPS):
    return configs.mfim(nx=16, ny=16, nz=36, lx=8e-9, ly=8e-9, nsteps=nsteps, seed=3)


def setup(text=None, **kw):
    kw.setdefault("mg_tol_rel", 1e-13)
    inp = fx.from_text(text or small_text(), **kw)
    sim = fx.FerroX(inp)
    # a polarized start (0.15 +- 0.05 C/m^2 in the FE cells), where every
    # Landau order contributes: from the +-1e-3 random start the gamma P^5
    # term is ~1e-10 of the others and a 1e-6 relative difference in gamma
    # moves the loss by its own rounding
    mask = sim.get("mask")
    rng0 = np.random.default_rng(5)
    P0 = np.where(mask == 0, 0.15 + 0.05 * rng0.uniform(-1, 1, mask.shape), 0.0)
    theta = sim.diff_get_params()
    V = triangular_sweep(NSTEPS, 3.0)
    rng = np.random.default_rng(11)
    target = rng.uniform(-0.05, 0.05, NSTEPS)
    return inp, sim, P0, theta, V, target


def loss_of(m, target):
    return float(np.mean((m - target) ** 2))


def dloss(m, target):
    return 2.0 * (m - target) / m.size


def adjoint(sim, P0, theta, V, target, K):
    sim.diff_set_params(theta)
    sim.set("Pz", P0)
    m = sim.diff_forward(V, K)
    gt, gv = sim.diff_backward(dloss(m, target))
    return m, gt, gv


def gpu_loss(sim, P0, theta, V, target):
    sim.diff_set_params(theta)
    sim.set("Pz", P0)
    return loss_of(sim.diff_forward(V, NSTEPS), target)


def test_adjoint_matches_central_differences_parameters():
    inp, sim, P0, theta, V, target = setup()
    m, gt, gv = adjoint(sim, P0, theta, V, target, K=7)
    for q in range(6):
        fd = central_fd(lambda t: gpu_loss(sim, P0, t, V, target), theta, q)
        err = abs(gt[q] - fd) / abs(fd)
        print(fx.FerroX.DIFF_PARAMS[q], gt[q], fd, err)
        assert err < FD_TOL, (q, gt[q], fd)


def test_adjoint_matches_central_differences_waveform():
    inp, sim, P0, theta, V, target = setup()
    m, gt, gv = adjoint(sim, P0, theta, V, target, K=7)

    def lv(v):
        return gpu_loss(sim, P0, theta, v, target)

    for n in (0, 9, 20, 33, NSTEPS - 1):
        fd = central_fd(lv, V, n, h_rel=1e-6)
        err = abs(gv[n] - fd) / max(abs(fd), 1e-12 * np.abs(gv).max())
        print("V", n, gv[n], fd, err)
        assert err < FD_TOL, (n, gv[n], fd)


def test_checkpoint_interval_does_not_change_the_gradient():
    inp, sim, P0, theta, V, target = setup()
    ref = None
    for K in (1, 7, NSTEPS):
        m, gt, gv = adjoint(sim, P0, theta, V, target, K)
        if ref is None:
            ref = (m, gt, gv)
            continue
        assert np.array_equal(m, ref[0])
        assert np.allclose(gt, ref[1], rtol=1e-12, atol=0)
        assert np.allclose(gv, ref[2], rtol=1e-12, atol=1e-12 * np.abs(ref[2]).max())


def test_forward_curve_matches_oracle_and_oracle_fd_gradient():
    """FP64 reference gradient: central differences of the CPU oracle's loss."""
    inp, sim, P0, theta, V, target = setup()
    m, gt, gv = adjoint(sim, P0, theta, V, target, K=10)

    def oracle_curve(t):
        kw = dict(zip(("alpha", "beta", "gamma", "g11", "g44", "epsilonX_fe"), t))
        kw["epsilonX_fe_tphase"] = t[5]
        o = Oracle(fx.from_text(small_text(), mg_tol_rel=1e-13, **kw))
        o.set("Pz", P0)
        o.set_bias_waveform(V)
        mask = o.get("mask")
        fe = mask == 0
        out = []
        for n in range(NSTEPS):
            o.step(1)
            out.append(o.get("Pz")[fe].mean())
        o.close()
        return np.array(out)

    mo = oracle_curve(theta)
    assert np.abs(m - mo).max() <= 1e-10 * np.abs(mo).max()
    for q in (0, 3, 5):
        fd = central_fd(lambda t: loss_of(oracle_curve(t), target), theta, q)
        err = abs(gt[q] - fd) / abs(fd)
        print("oracle", fx.FerroX.DIFF_PARAMS[q], gt[q], fd, err)
        assert err < FD_TOL


def test_torch_autograd_function():
    import torch
    inp, sim, P0, theta, V, target = setup()
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    Vt = torch.tensor(V, dtype=torch.float64, requires_grad=True)
    tg = torch.tensor(target, dtype=torch.float64)
    m = pv_curve(sim, th, Vt, P0, K=8)
    loss = torch.mean((m - tg) ** 2)
    loss.backward()
    m2, gt, gv = adjoint(sim, P0, theta, V, target, K=8)
    assert np.allclose(th.grad.numpy(), gt, rtol=1e-12, atol=0)
    assert np.allclose(Vt.grad.numpy(), gv, rtol=1e-12, atol=1e-14 * np.abs(gv).max())
    assert abs(loss.item() - loss_of(m2, target)) <= 1e-14 * loss.item()


def test_gradient_descent_fits_the_pv_curve():
    """Fitting: a target curve from perturbed parameters; a few Adam steps on
    log|theta| driven by the adjoint gradient reduce the misfit."""
    import torch
    inp, sim, P0, theta, V, _ = setup()
    rng = np.random.default_rng(7)
    theta_star = theta * (1 + rng.uniform(-0.1, 0.1, 6))
    sim.diff_set_params(theta_star)
    sim.set("Pz", P0)
    target = torch.tensor(sim.diff_forward(V, 10))
    sign = torch.tensor(np.sign(theta))
    logt = torch.tensor(np.log(np.abs(theta)), requires_grad=True)
    Vt = torch.tensor(V)
    opt = torch.optim.Adam([logt], lr=0.02)
    losses = []
    for it in range(15):
        opt.zero_grad()
        m = pv_curve(sim, sign * torch.exp(logt), Vt, P0, K=10)
        loss = torch.mean((m - target) ** 2) / torch.mean(target ** 2)
        loss.backward()
        opt.step()
        losses.append(loss.item())
    print("relative misfit", losses[0], "->", losses[-1])
    assert losses[-1] < 0.2 * losses[0]
