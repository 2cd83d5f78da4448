"""Differentiation form: the P-V curve of the forward-Euler TDGL + Poisson
time loop as a torch.autograd.Function over the C ABI.

forward(theta, V) runs the device time loop (fx_diff_forward) from a fixed
initial polarization with the gate bias V[n] at step n and returns
m[n] = mean Pz over the ferroelectric cells after step n.  backward(dL/dm)
runs the hand-written discrete adjoint (fx_diff_backward: checkpoints every
K steps, segment recompute, adjoint Poisson solves on the same multigrid) and
returns dL/dtheta and dL/dV.  theta = (alpha, beta, gamma, g11, g44,
epsilonX_fe) in the inputs' units; everything is FP64.

The reference has no counterpart: its parameters are process-global inputs
read once (FerroX.cpp:327-472) and its gate bias changes only at sweep events
(main.cpp:572-604).
"""
from __future__ import annotations

import numpy as np
import torch

from .solver import FerroX

PARAMS = FerroX.DIFF_PARAMS


class PVCurve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, V: torch.Tensor, sim: FerroX, P0: np.ndarray, K: int) -> torch.Tensor:
        sim.diff_set_params(theta.detach().cpu().to(torch.float64).numpy())
        sim.set("Pz", P0)
        m = sim.diff_forward(V.detach().cpu().to(torch.float64).numpy(), K)
        ctx.sim = sim
        return torch.from_numpy(m)

    @staticmethod
    def backward(ctx, gm: torch.Tensor):
        gt, gv = ctx.sim.diff_backward(gm.detach().cpu().to(torch.float64).contiguous().numpy())
        return torch.from_numpy(gt), torch.from_numpy(gv), None, None, None


def pv_curve(sim: FerroX, theta: torch.Tensor, V: torch.Tensor, P0: np.ndarray, K: int = 20) -> torch.Tensor:
    """Differentiable P-V curve (see the module docstring)."""
    return PVCurve.apply(theta, V, sim, P0, K)


def triangular_sweep(nsteps: int, vmax: float) -> np.ndarray:
    """0 -> +vmax -> -vmax -> 0, linear, nsteps values (the bias of steps 1..nsteps)."""
    t = (np.arange(1, nsteps + 1) / nsteps) * 4.0      # quarter periods
    return vmax * np.where(t <= 1, t, np.where(t <= 3, 2 - t, t - 4))


def central_fd(f, x: np.ndarray, idx, h_rel: float = 1e-6) -> float:
    """Central difference of the scalar f at x along coordinate idx with step
    h_rel * |x[idx]| (or h_rel when x[idx] is 0)."""
    h = h_rel * abs(x[idx]) if x[idx] != 0 else h_rel
    xp = x.copy(); xp[idx] += h
    xm = x.copy(); xm[idx] -= h
    return (f(xp) - f(xm)) / (2 * h)
