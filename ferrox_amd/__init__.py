"""ferrox_amd -- MI355X-native FerroX TDGL + Poisson hot path.

The compute path is libferrox_mi.so (HIP kernels for gfx950 behind the C ABI
in include/ferrox_mi.h).  This package is the Python host mirror of the
reference's interface for that path; it has no CPU fallback.
"""
from ._abi import FIELDS, FIELD_ID, PLOT_VARS, FerroXError, FxInputs, load_library
from .inputs import apply, default_inputs, finalize, from_text, parse_file
from .solver import FerroX

__all__ = ["FIELDS", "FIELD_ID", "PLOT_VARS", "FerroXError", "FxInputs", "load_library", "apply",
           "default_inputs", "finalize", "from_text", "parse_file", "FerroX"]
