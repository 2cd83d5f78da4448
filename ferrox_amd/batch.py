"""Batching form: many independent MFIM devices advanced together on one GPU
(libferrox_mi's fx_batch_*: one persistent launch per stage over all devices,
captured step graphs, autotuned launch shapes, a direct spectral Poisson solve
for laterally periodic layered stacks).  No reference counterpart: the
reference runs one device per process (main.cpp:48-623).
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from ._abi import FIELD_ID, FerroXError, FxInputs, load_library


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Batch:
    def __init__(self, inputs: list[FxInputs]):
        self._lib = load_library()
        self.inputs = list(inputs)
        arr = (C.POINTER(FxInputs) * len(self.inputs))(*[C.pointer(i) for i in self.inputs])
        h = C.c_void_p()
        rc = self._lib.fx_batch_create(len(self.inputs), arr, C.byref(h))
        if rc:
            raise FerroXError(rc, "fx_batch_create failed (see stderr)")
        self._h = h
        self.shapes = []
        d = (C.c_int * 3)()
        for k in range(len(self.inputs)):
            self._check(self._lib.fx_batch_dims(self._h, k, d))
            self.shapes.append((d[2], d[1], d[0]))

    def _check(self, rc):
        if rc:
            raise FerroXError(rc, (self._lib.fx_batch_last_error(self._h) or b"").decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fx_batch_destroy(self._h)
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

    @property
    def num_devices(self) -> int:
        return len(self.shapes)

    @property
    def cells(self) -> int:
        return int(self._lib.fx_batch_cells(self._h))

    def step(self, nsteps: int, graph: bool = True) -> float:
        """nsteps steps of every device; returns the wall time in seconds."""
        w = C.c_double()
        self._check(self._lib.fx_batch_step(self._h, int(nsteps), 1 if graph else 0, C.byref(w)))
        return w.value

    def get(self, dev: int, name: str) -> np.ndarray:
        out = np.empty(self.shapes[dev])
        self._check(self._lib.fx_batch_get_field(self._h, dev, FIELD_ID[name], _dp(out)))
        return out

    def set(self, dev: int, name: str, value):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=np.float64), self.shapes[dev]))
        self._check(self._lib.fx_batch_set_field(self._h, dev, FIELD_ID[name], _dp(a)))

    def autotune(self) -> dict:
        buf = C.create_string_buffer(1 << 16)
        self._check(self._lib.fx_batch_autotune(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())


def basis(n: int) -> tuple[np.ndarray, np.ndarray]:
    """The library's real Fourier basis of an n-point periodic direction."""
    lib = load_library()
    Q = np.empty((n, n))
    mu = np.empty(n)
    rc = lib.fx_batch_basis(n, _dp(Q), _dp(mu))
    if rc:
        raise FerroXError(rc, "fx_batch_basis")
    return Q, mu
