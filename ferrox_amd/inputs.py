"""FerroX inputs: parse inputs files / key=value overrides into fx_inputs.

The reader itself is the library's C++ ParmParse-compatible parser
(fx_inputs_parse_file / fx_inputs_set), so the Python caller, the C++ command
line driver and the tests see identical key handling.  Key names and defaults
are the reference's (Source/FerroX.cpp:273-572).
"""
from __future__ import annotations

import ctypes as C

from ._abi import FxInputs, FerroXError, load_library


def default_inputs() -> FxInputs:
    lib = load_library()
    inp = FxInputs()
    rc = lib.fx_inputs_default(C.byref(inp))
    if rc:
        raise FerroXError(rc, "fx_inputs_default failed")
    return inp


def apply(inp: FxInputs, text: str) -> FxInputs:
    """Apply inputs-file text (one or more 'key = value' lines)."""
    lib = load_library()
    err = C.create_string_buffer(512)
    rc = lib.fx_inputs_set(C.byref(inp), text.encode(), err, 512)
    if rc:
        raise FerroXError(rc, err.value.decode())
    return inp


def parse_file(path: str, inp: FxInputs | None = None) -> FxInputs:
    lib = load_library()
    inp = inp if inp is not None else default_inputs()
    err = C.create_string_buffer(512)
    rc = lib.fx_inputs_parse_file(str(path).encode(), C.byref(inp), err, 512)
    if rc:
        raise FerroXError(rc, err.value.decode())
    return inp


def finalize(inp: FxInputs) -> FxInputs:
    lib = load_library()
    err = C.create_string_buffer(512)
    rc = lib.fx_inputs_finalize(C.byref(inp), err, 512)
    if rc:
        raise FerroXError(rc, err.value.decode())
    return inp


def from_text(text: str, **overrides) -> FxInputs:
    """Defaults + inputs text + keyword overrides, finalized."""
    inp = default_inputs()
    apply(inp, text)
    if overrides:
        apply(inp, "\n".join(f"{k} = {_fmt(v)}" for k, v in overrides.items()))
    return finalize(inp)


def _fmt(v) -> str:
    import numbers
    if isinstance(v, (list, tuple)) or (hasattr(v, "ndim") and getattr(v, "ndim", 0) == 1):
        return " ".join(_fmt(x) for x in v)
    if isinstance(v, bool):
        return str(int(v))
    if isinstance(v, numbers.Integral):
        return str(int(v))
    if isinstance(v, numbers.Real):
        return repr(float(v))        # shortest round-trip text, numpy scalars included
    return str(v)
