"""Python mirror of the reference's hot-path interface over the C ABI.

Method names follow the reference functions they replace
(Source/Solver/ElectrostaticSolver.H, TotalEnergyDensity.H, ChargeDensity.H,
Source/main.cpp's time loop).  Arrays are numpy float64 of shape (nz, ny, nx):
memory order i fastest, the order of the reference's single-box FABs.
All arithmetic runs in libferrox_mi.so on the GPU; this module only moves
buffers and status codes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import (FIELD_ID, FxInputs, FxSolveStats, FxStepStats, FerroXError,
                   load_library, stats_dict)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class FerroX:
    """One simulated device (one box of cells) on one GPU, one HIP stream."""

    def __init__(self, inputs: FxInputs):
        self._lib = load_library()
        self.inputs = inputs
        h = C.c_void_p()
        rc = self._lib.fx_create(C.byref(inputs), C.byref(h))
        if rc:
            raise FerroXError(rc, "fx_create failed (see stderr)")
        self._h = h
        dims = (C.c_int * 3)()
        self._check(self._lib.fx_get_dims(self._h, dims))
        self.n = (dims[0], dims[1], dims[2])
        self.shape = (dims[2], dims[1], dims[0])

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc: int):
        if rc:
            msg = self._lib.fx_last_error(self._h) if getattr(self, "_h", None) else b""
            raise FerroXError(rc, (msg or b"").decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- fields -----------------------------------------------------------
    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        self._check(self._lib.fx_get_field(self._h, FIELD_ID[name], _dp(out)))
        return out

    def set(self, name: str, value):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=np.float64), self.shape))
        self._check(self._lib.fx_set_field(self._h, FIELD_ID[name], _dp(a)))

    def device_ptr(self, name: str) -> int:
        p = C.c_void_p()
        self._check(self._lib.fx_field_device_ptr(self._h, FIELD_ID[name], C.byref(p)))
        return p.value

    @property
    def stream(self) -> int:
        return self._lib.fx_stream(self._h)

    def sync(self):
        self._check(self._lib.fx_sync(self._h))

    # -- hot path (reference names) ----------------------------------------
    def SetPhiBC(self, phi_bc_lo: float, phi_bc_hi: float):
        """SetPhiBC_z + setLevelBC (ElectrostaticSolver.cpp:502-524)."""
        self._check(self._lib.fx_set_bias(self._h, phi_bc_lo, phi_bc_hi))

    def set_bias_waveform(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self._lib.fx_set_bias_waveform(self._h, _dp(v), int(v.size)))

    def ComputePhi_Rho(self) -> dict:
        """ComputePhi_Rho (ElectrostaticSolver.cpp:678-747)."""
        st = FxSolveStats()
        self._check(self._lib.fx_compute_phi_rho(self._h, C.byref(st)))
        return stats_dict(st)

    def ComputeEfromPhi(self):
        """ComputeEfromPhi (ElectrostaticSolver.cpp:134-203)."""
        self._check(self._lib.fx_compute_E(self._h))

    def CalculateTDGL_RHS(self):
        """Calculate_Landau/_Grad/_Elec + CalculateTDGL_RHS (TotalEnergyDensity.cpp)."""
        self._check(self._lib.fx_tdgl_rhs(self._h))

    def update(self, order: int = 1):
        """P update of main.cpp:293-348 (order 1 forward Euler, 2 Heun)."""
        self._check(self._lib.fx_update(self._h, order))

    def step(self, nsteps: int = 0) -> dict:
        """nsteps iterations of main.cpp:261-623."""
        st = FxStepStats()
        self._check(self._lib.fx_step(self._h, nsteps, C.byref(st)))
        return stats_dict(st)

    def poisson_solve(self, rhs, acoef=None, phi0=None) -> tuple[np.ndarray, dict]:
        """MLMG::solve replacement on this context's geometry and bias."""
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(rhs, dtype=np.float64), self.shape))
        phi = np.zeros(self.shape) if phi0 is None else np.array(phi0, dtype=np.float64, copy=True)
        a = None if acoef is None else np.ascontiguousarray(np.broadcast_to(np.asarray(acoef, dtype=np.float64), self.shape))
        st = FxSolveStats()
        rc = self._lib.fx_poisson_solve(self._h, _dp(b), None if a is None else _dp(a), _dp(phi), C.byref(st))
        self._check(rc)
        return phi, stats_dict(st)

    # -- plotfiles -----------------------------------------------------------
    def write_plotfile(self, directory: str, time: float = 0.0, step: int = 0):
        """AMReX plotfile of the current state (WritePlotfile, Plotfile.cpp:5-177)."""
        self._check(self._lib.fx_write_plotfile(self._h, str(directory).encode(), float(time), int(step)))

    def set_plot_dir(self, directory: str):
        """Where fx_step puts its plt%08d directories when plot_int > 0."""
        self._check(self._lib.fx_set_plot_dir(self._h, str(directory).encode()))

    # -- differentiation form (extension) ------------------------------------
    DIFF_PARAMS = ("alpha", "beta", "gamma", "g11", "g44", "epsilonX_fe")

    def diff_set_params(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float64)
        assert t.shape == (6,)
        self._check(self._lib.fx_diff_set_params(self._h, _dp(t)))

    def diff_get_params(self) -> np.ndarray:
        t = np.empty(6)
        self._check(self._lib.fx_diff_get_params(self._h, _dp(t)))
        return t

    def diff_forward(self, V, K: int = 20) -> np.ndarray:
        """Euler steps from the current state with top bias V[n] at step n;
        returns the mean Pz over the ferroelectric cells after each step."""
        v = np.ascontiguousarray(V, dtype=np.float64)
        m = np.empty(v.size)
        self._check(self._lib.fx_diff_forward(self._h, _dp(v), int(v.size), int(K), _dp(m)))
        return m

    def diff_backward(self, dL_dm) -> tuple[np.ndarray, np.ndarray]:
        """(dL/dtheta, dL/dV) through the last diff_forward."""
        g = np.ascontiguousarray(dL_dm, dtype=np.float64)
        gt = np.empty(6)
        gv = np.empty(g.size)
        self._check(self._lib.fx_diff_backward(self._h, _dp(g), _dp(gt), _dp(gv)))
        return gt, gv

    def bench_kernel(self, which: str, reps: int = 20) -> float:
        """Average ms per launch of a hot-path kernel (HIP events, context stream)."""
        ms = C.c_double()
        self._check(self._lib.fx_bench_kernel(self._h, which.encode(), reps, C.byref(ms)))
        return ms.value

    def levels(self) -> list[tuple[int, int, int]]:
        out = []
        d = (C.c_int * 3)()
        for l in range(self._lib.fx_num_levels(self._h)):
            self._check(self._lib.fx_level_dims(self._h, l, d))
            out.append((d[0], d[1], d[2]))
        return out
