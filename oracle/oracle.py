"""ORACLE -- test infrastructure only.

ctypes wrapper of oracle/_build/libferrox_oracle.so, the FP64 CPU restatement
of FerroX's hot path (see ferrox_oracle.cpp for the reference lines each part
follows).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may use it; the product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "libferrox_oracle.so")


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        from ferrox_amd._abi import FxSolveStats, FxStepStats, FxInputs  # struct layouts only
        L = C.CDLL(LIB)
        P, DP = C.c_void_p, C.POINTER(C.c_double)
        sig = {
            "fxo_create": (P, [C.POINTER(FxInputs)]),
            "fxo_destroy": (None, [P]),
            "fxo_error": (C.c_char_p, [P]),
            "fxo_num_levels": (C.c_int, [P]),
            "fxo_level_dims": (C.c_int, [P, C.c_int, C.POINTER(C.c_int)]),
            "fxo_get_field": (C.c_int, [P, C.c_int, DP]),
            "fxo_set_field": (C.c_int, [P, C.c_int, DP]),
            "fxo_set_bias": (C.c_int, [P, C.c_double, C.c_double]),
            "fxo_set_bias_waveform": (C.c_int, [P, DP, C.c_int]),
            "fxo_compute_phi_rho": (C.c_int, [P, C.POINTER(FxSolveStats)]),
            "fxo_compute_E": (C.c_int, [P]),
            "fxo_compute_rho": (C.c_int, [P]),
            "fxo_tdgl_rhs": (C.c_int, [P]),
            "fxo_update": (C.c_int, [P, C.c_int]),
            "fxo_step": (C.c_int, [P, C.c_int, C.POINTER(FxStepStats)]),
            "fxo_poisson_solve": (C.c_int, [P, DP, DP, DP, C.c_double, C.POINTER(FxSolveStats)]),
            "fxo_apply_op": (C.c_int, [P, DP, DP, DP, C.c_int]),
            "fxo_get_faces": (C.c_int, [P, C.c_int, DP]),
            "fxo_threads": (C.c_int, []),
            "fxo_set_threads": (C.c_int, [C.c_int]),
            "fxo_uniform01": (C.c_double, [C.c_ulonglong, C.c_ulonglong]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleError(RuntimeError):
    pass


class Oracle:
    """Same surface as ferrox_amd.FerroX, computed by the CPU restatement."""

    def __init__(self, inputs):
        from ferrox_amd._abi import FIELD_ID
        self._fid = FIELD_ID
        self._L = lib()
        self.inputs = inputs
        self._h = self._L.fxo_create(C.byref(inputs))
        n = inputs.n_cell
        self.n = (n[0], n[1], n[2])
        self.shape = (n[2], n[1], n[0])

    def close(self):
        if self._h:
            self._L.fxo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise OracleError(f"oracle status {rc}: {self._L.fxo_error(self._h).decode()}")

    def get(self, name):
        out = np.empty(self.shape)
        self._check(self._L.fxo_get_field(self._h, self._fid[name], _dp(out)))
        return out

    def set(self, name, value):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=np.float64), self.shape))
        self._check(self._L.fxo_set_field(self._h, self._fid[name], _dp(a)))

    def SetPhiBC(self, lo, hi):
        self._check(self._L.fxo_set_bias(self._h, lo, hi))

    def set_bias_waveform(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self._L.fxo_set_bias_waveform(self._h, _dp(v), int(v.size)))

    def ComputePhi_Rho(self):
        from ferrox_amd._abi import FxSolveStats, stats_dict
        st = FxSolveStats()
        self._check(self._L.fxo_compute_phi_rho(self._h, C.byref(st)))
        return stats_dict(st)

    def ComputeEfromPhi(self):
        self._check(self._L.fxo_compute_E(self._h))

    def ComputeRho(self):
        """rho, electrons and holes of the current phi (no solve)."""
        self._check(self._L.fxo_compute_rho(self._h))

    def CalculateTDGL_RHS(self):
        self._check(self._L.fxo_tdgl_rhs(self._h))

    def update(self, order=1):
        self._check(self._L.fxo_update(self._h, order))

    def step(self, nsteps=0):
        from ferrox_amd._abi import FxStepStats, stats_dict
        st = FxStepStats()
        self._check(self._L.fxo_step(self._h, nsteps, C.byref(st)))
        return stats_dict(st)

    def poisson_solve(self, rhs, acoef=None, phi0=None, tol=None):
        from ferrox_amd._abi import FxSolveStats, stats_dict
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(rhs, dtype=np.float64), self.shape))
        phi = np.zeros(self.shape) if phi0 is None else np.array(phi0, dtype=np.float64, copy=True)
        a = None if acoef is None else np.ascontiguousarray(np.broadcast_to(np.asarray(acoef, dtype=np.float64), self.shape))
        st = FxSolveStats()
        t = self.inputs.mg_tol_rel if tol is None else tol
        self._check(self._L.fxo_poisson_solve(self._h, _dp(b), None if a is None else _dp(a), _dp(phi), t, C.byref(st)))
        return phi, stats_dict(st)

    def apply_op(self, x, acoef=None, inhomog=False):
        xx = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.shape)
        a = None if acoef is None else np.ascontiguousarray(np.broadcast_to(np.asarray(acoef, dtype=np.float64), self.shape))
        self._check(self._L.fxo_apply_op(self._h, None if a is None else _dp(a), _dp(xx), _dp(y), int(inhomog)))
        return y

    def faces(self):
        nx, ny, nz = self.n
        out = []
        for w, shp in enumerate([(nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)]):
            a = np.empty(shp)
            self._check(self._L.fxo_get_faces(self._h, w, _dp(a)))
            out.append(a)
        return out

    def levels(self):
        d = (C.c_int * 3)()
        out = []
        for l in range(self._L.fxo_num_levels(self._h)):
            self._L.fxo_level_dims(self._h, l, d)
            out.append((d[0], d[1], d[2]))
        return out


def threads() -> int:
    return lib().fxo_threads()


def set_threads(n: int) -> int:
    """Threads of the oracle's parallel loops from now on; returns the count in effect."""
    return lib().fxo_set_threads(int(n))


def uniform01(seed: int, idx: int) -> float:
    return lib().fxo_uniform01(seed, idx)
