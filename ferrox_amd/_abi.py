"""ctypes mirror of include/ferrox_mi.h (the C ABI of libferrox_mi.so).

The structures below must match the header field for field; tests check
sizeof(fx_inputs) against the library's own view through fx_inputs_default.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libferrox_mi.so")

FX_OK, FX_EINVAL, FX_ENOCONV_MG, FX_ENOCONV_OUTER, FX_ENAN, FX_EHIP, FX_ENOMEM = range(7)
STATUS_NAMES = {0: "FX_OK", 1: "FX_EINVAL", 2: "FX_ENOCONV_MG", 3: "FX_ENOCONV_OUTER",
                4: "FX_ENAN", 5: "FX_EHIP", 6: "FX_ENOMEM"}
FX_BC_PERIODIC, FX_BC_DIRICHLET, FX_BC_NEUMANN = 0, 1, 2
FX_GHOST_AMREX, FX_GHOST_FACE = 0, 1

FIELDS = ["Px", "Py", "Pz", "Phi", "PoissonRHS", "Ex", "Ey", "Ez", "holes", "electrons",
          "charge", "epsilon", "mask", "tphase", "alpha", "beta", "theta", "PhiDiff",
          "alpha_cc", "Gamma", "GL_rhs_x", "GL_rhs_y", "GL_rhs_z", "landau_alpha"]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}
# plotfile variable order of the reference (Plotfile.cpp:30-106)
PLOT_VARS = FIELDS[:18]

_D3 = C.c_double * 3
_I3 = C.c_int * 3


class FxInputs(C.Structure):
    _fields_ = [
        ("n_cell", _I3), ("prob_lo", _D3), ("prob_hi", _D3), ("is_periodic", _I3),
        ("max_grid_size", _I3), ("bc_lo", _I3), ("bc_hi", _I3),
        ("TimeIntegratorOrder", C.c_int), ("prob_type", C.c_int), ("is_polarization_scalar", C.c_int),
        ("nsteps", C.c_int), ("plot_int", C.c_int), ("dt", C.c_double), ("delta", C.c_double),
        ("random_seed", C.c_int), ("voltage_sweep", C.c_int), ("inc_step", C.c_int), ("num_Vapp_max", C.c_int),
        ("Phi_Bc_lo", C.c_double), ("Phi_Bc_hi", C.c_double), ("Phi_Bc_inc", C.c_double),
        ("Phi_Bc_hi_max", C.c_double), ("phi_tolerance", C.c_double),
        ("include_Landau", C.c_int), ("include_Grad", C.c_int), ("include_Elec", C.c_int),
        ("P_BC_flag_lo", _I3), ("P_BC_flag_hi", _I3), ("lambda_", C.c_double),
        ("epsilon_0", C.c_double), ("epsilonX_fe", C.c_double), ("epsilonX_fe_tphase", C.c_double),
        ("epsilonZ_fe", C.c_double), ("epsilon_de", C.c_double), ("epsilon_si", C.c_double),
        ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double),
        ("alpha_12", C.c_double), ("alpha_112", C.c_double), ("alpha_123", C.c_double),
        ("BigGamma", C.c_double), ("g11", C.c_double), ("g44", C.c_double), ("g12", C.c_double), ("g44_p", C.c_double),
        ("FE_lo", _D3), ("FE_hi", _D3), ("DE_lo", _D3), ("DE_hi", _D3), ("SC_lo", _D3), ("SC_hi", _D3),
        ("Channel_lo", _D3), ("Channel_hi", _D3), ("Remnant_P", _D3),
        ("acceptor_doping", C.c_double), ("donor_doping", C.c_double),
        ("use_Fermi_Dirac", C.c_int), ("use_work_function", C.c_int), ("metal_work_function", C.c_double),
        ("Coordinate_Transformation", C.c_int), ("use_Euler_angles", C.c_int),
        ("mg_tol_rel", C.c_double), ("mg_max_iter", C.c_int), ("mg_nu1", C.c_int), ("mg_nu2", C.c_int),
        ("mg_warm_start", C.c_int), ("outer_max_iter", C.c_int), ("phi_ghost_mode", C.c_int),
        ("precision", C.c_int), ("bias_ramp_steps", C.c_int), ("bias_ramp_start", C.c_double),
        ("bias_ramp_end", C.c_double), ("verbose", C.c_int), ("mg_fused_sweep", C.c_int),
        ("mg_block_levels", C.c_int), ("mg_stream_min_cells", C.c_int), ("mg_coarsen_ratio", C.c_double), ("mg_exact_tail", C.c_int),
        ("mg_refine_nu1", C.c_int), ("mg_refine_nu2", C.c_int),
        ("poisson_direct", C.c_int), ("poisson_direct_check", C.c_int), ("tdgl_fused_rhs", C.c_int), ("tdgl_stream", C.c_int),
        ("poisson_pcg", C.c_int),
        ("plot_Phi", C.c_int), ("plot_PoissonRHS", C.c_int), ("plot_E", C.c_int), ("plot_holes", C.c_int), ("plot_electrons", C.c_int), ("plot_charge", C.c_int), ("plot_epsilon", C.c_int), ("plot_mask", C.c_int), ("plot_tphase", C.c_int), ("plot_alpha", C.c_int), ("plot_beta", C.c_int), ("plot_theta", C.c_int), ("plot_PhiDiff", C.c_int),
        ("tphase_geom_function", C.c_char * 4096), ("angle_alpha_function", C.c_char * 2048),
        ("angle_beta_function", C.c_char * 2048), ("angle_theta_function", C.c_char * 2048),
    ]


class FxSolveStats(C.Structure):
    _fields_ = [("outer_iters", C.c_int), ("vcycles", C.c_int), ("outer_err", C.c_double),
                ("mg_resid0", C.c_double), ("mg_resid", C.c_double), ("mg_bnorm", C.c_double)]


class FxStepStats(C.Structure):
    _fields_ = [("steps_done", C.c_int), ("stopped_steady", C.c_int), ("total_vcycles", C.c_int),
                ("total_outer_iters", C.c_int), ("total_solves", C.c_int), ("last_phi_change", C.c_double),
                ("time", C.c_double), ("Phi_Bc_hi", C.c_double), ("num_Vapp", C.c_int),
                ("wall_seconds", C.c_double)]


def stats_dict(s) -> dict:
    return {name: getattr(s, name) for name, _ in s._fields_}


# Every entry point the header declares, with (restype, argtypes).
_P = C.c_void_p
_DP = C.POINTER(C.c_double)
SIGNATURES = {
    "fx_version": (C.c_char_p, []),
    "fx_inputs_size": (C.c_int, []),
    "fx_inputs_default": (C.c_int, [C.POINTER(FxInputs)]),
    "fx_inputs_parse_file": (C.c_int, [C.c_char_p, C.POINTER(FxInputs), C.c_char_p, C.c_int]),
    "fx_inputs_set": (C.c_int, [C.POINTER(FxInputs), C.c_char_p, C.c_char_p, C.c_int]),
    "fx_inputs_finalize": (C.c_int, [C.POINTER(FxInputs), C.c_char_p, C.c_int]),
    "fx_create": (C.c_int, [C.POINTER(FxInputs), C.POINTER(_P)]),
    "fx_destroy": (C.c_int, [_P]),
    "fx_last_error": (C.c_char_p, [_P]),
    "fx_get_dims": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "fx_get_field": (C.c_int, [_P, C.c_int, _DP]),
    "fx_set_field": (C.c_int, [_P, C.c_int, _DP]),
    "fx_set_bias": (C.c_int, [_P, C.c_double, C.c_double]),
    "fx_set_bias_waveform": (C.c_int, [_P, _DP, C.c_int]),
    "fx_diff_set_params": (C.c_int, [_P, _DP]),
    "fx_eval_cell_function": (C.c_int, [C.POINTER(FxInputs), C.c_char_p, _DP, C.c_char_p, C.c_int]),
    "fx_write_plotfile": (C.c_int, [_P, C.c_char_p, C.c_double, C.c_int]),
    "fx_set_plot_dir": (C.c_int, [_P, C.c_char_p]),
    "fx_batch_create": (C.c_int, [C.c_int, C.POINTER(C.POINTER(FxInputs)), C.POINTER(_P)]),
    "fx_batch_destroy": (C.c_int, [_P]),
    "fx_batch_last_error": (C.c_char_p, [_P]),
    "fx_batch_num_devices": (C.c_int, [_P]),
    "fx_batch_cells": (C.c_longlong, [_P]),
    "fx_batch_dims": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    "fx_batch_step": (C.c_int, [_P, C.c_int, C.c_int, _DP]),
    "fx_batch_get_field": (C.c_int, [_P, C.c_int, C.c_int, _DP]),
    "fx_batch_set_field": (C.c_int, [_P, C.c_int, C.c_int, _DP]),
    "fx_batch_autotune": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "fx_batch_basis": (C.c_int, [C.c_int, _DP, _DP]),
    "fx_diff_get_params": (C.c_int, [_P, _DP]),
    "fx_diff_forward": (C.c_int, [_P, _DP, C.c_int, C.c_int, _DP]),
    "fx_diff_backward": (C.c_int, [_P, _DP, _DP, _DP]),
    "fx_compute_phi_rho": (C.c_int, [_P, C.POINTER(FxSolveStats)]),
    "fx_compute_E": (C.c_int, [_P]),
    "fx_tdgl_rhs": (C.c_int, [_P]),
    "fx_update": (C.c_int, [_P, C.c_int]),
    "fx_step": (C.c_int, [_P, C.c_int, C.POINTER(FxStepStats)]),
    "fx_poisson_solve": (C.c_int, [_P, _DP, _DP, _DP, C.POINTER(FxSolveStats)]),
    "fx_poisson_solve_noalloc": (C.c_int, [_P, C.c_void_p, C.c_void_p]),
    "fx_field_device_ptr": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "fx_sync": (C.c_int, [_P]),
    "fx_stream": (_P, [_P]),
    "fx_bench_kernel": (C.c_int, [_P, C.c_char_p, C.c_int, C.POINTER(C.c_double)]),
    "fx_num_levels": (C.c_int, [_P]),
    "fx_level_dims": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
}


class FerroXError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load_library(path: str | None = None):
    """Load libferrox_mi.so.  There is no fallback: a missing or stale build
    raises, so the hot path can never silently run anywhere but the HIP
    kernels."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
