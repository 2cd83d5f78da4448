"""C-ABI library and inputs reader (no GPU needed: nothing here launches a kernel)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import ferrox_amd as fx
from ferrox_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ferrox_mi.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_abi.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the ctypes mirror binds every one of them
    assert set(syms) <= set(_abi.SIGNATURES), set(syms) - set(_abi.SIGNATURES)


def test_struct_layout_matches_library():
    lib = fx.load_library()
    assert lib.fx_inputs_size() == C.sizeof(_abi.FxInputs)
    assert lib.fx_version().decode().startswith("ferrox")


REF_LIKE = """
#################################
domain.prob_lo = -16.e-9 -16.e-9 0.e-9
domain.prob_hi =  16.e-9  16.e-9 9.e-9
domain.n_cell = 64 64 18
domain.max_grid_size = 64 64 18
domain.coord_sys = cartesian
prob_type = 1
TimeIntegratorOrder = 1
nsteps = 1000
plot_int = 100
dt = 2.0e-13
P_BC_flag_lo = 3 3 0
P_BC_flag_hi = 3 3 1
lambda = 3.0e-9
domain.is_periodic = 1 1 0
boundary.hi = per per dir(0.0)
boundary.lo = per per dir(0.0)
Phi_Bc_lo = 0.0
Phi_Bc_hi = 0.0
inc_step = 5000
Phi_Bc_inc = 0.0
SC_lo = -1.0 -1.0 -1.0
SC_hi = -1.0 -1.0 -1.0
DE_lo = -16.e-9 -16.e-9 0.0e-9
DE_hi =  16.e-9  16.e-9 4.0e-9
FE_lo = -16.e-9 -16.e-9 4.0e-9
FE_hi =  16.e-9  16.e-9 9.e-9
epsilon_0 = 8.85e-12
epsilonX_fe = 24.0
epsilonZ_fe = 24.0
epsilon_de = 10.0
epsilon_si = 11.7
alpha = -2.5e9
beta = 6.0e10
gamma = 1.5e11
BigGamma = 100
g11 = 1.0e-9
g44 = 1.0e-9
g44_p = 0.0
g12 = 0.0
alpha_12 = 0.0
alpha_112 = 0.0
alpha_123 = 0.0
amrex.the_arena_is_managed = 1   # unknown keys are ignored
"""


def test_reference_style_inputs_file(tmp_path):
    p = tmp_path / "inputs"
    p.write_text(REF_LIKE)
    inp = fx.finalize(fx.parse_file(str(p)))
    assert list(inp.n_cell) == [64, 64, 18]
    assert inp.prob_lo[0] == -16e-9 and inp.prob_hi[2] == 9e-9
    assert list(inp.is_periodic) == [1, 1, 0]
    assert list(inp.bc_lo) == [0, 0, 1] and list(inp.bc_hi) == [0, 0, 1]
    assert list(inp.P_BC_flag_lo) == [3, 3, 0] and list(inp.P_BC_flag_hi) == [3, 3, 1]
    assert inp.lambda_ == 3e-9 and inp.dt == 2e-13 and inp.nsteps == 1000 and inp.inc_step == 5000
    assert inp.alpha == -2.5e9 and inp.g11 == 1e-9
    # defaults of the reference (FerroX.cpp:273-572)
    assert inp.is_polarization_scalar == 1 and inp.use_Fermi_Dirac == 1 and inp.delta == 1e-6
    assert inp.phi_tolerance == 1e-7 and inp.random_seed == 1 and inp.num_Vapp_max == 1
    assert inp.epsilonX_fe_tphase == 24.0
    assert tuple(inp.Remnant_P) == (0.0, 0.0, 0.002)
    assert inp.Channel_lo[0] == -16e-9 - 1.0 and inp.Channel_hi[2] == 9e-9 + 1.0
    assert inp.mg_tol_rel == 1e-10


def test_last_definition_wins_and_expressions():
    inp = fx.default_inputs()
    fx.apply(inp, REF_LIKE + "\nnsteps = 7\nmy_constants.L = 16.e-9\nmy_constants.nz = 9\n"
             "domain.prob_hi = L L nz*1.0e-9\ndt = 2**-3*1.0e-12\n")
    fx.finalize(inp)
    assert inp.nsteps == 7
    assert inp.prob_hi[0] == 16e-9 and abs(inp.prob_hi[2] - 9e-9) < 1e-20
    assert inp.dt == 0.125e-12


@pytest.mark.parametrize("bad,msg", [
    ("", "missing required key"),
    ("P_BC_flag_lo = 3 3 7", "P_BC_flag"),
    ("prob_type = 4", "prob_type"),
    ("boundary.lo = per dir(0) dir(0)", "per"),
    ("use_sundials = 1", "SUNDIALS"),
])
def test_input_errors(bad, msg):
    inp = fx.default_inputs()
    text = REF_LIKE if bad else REF_LIKE.replace("alpha = -2.5e9", "")
    with pytest.raises(fx.FerroXError) as e:
        fx.apply(inp, text + "\n" + bad)
        fx.finalize(inp)
    assert msg in str(e.value)


def test_cli_driver_is_built():
    exe = os.path.join(ROOT, "ferrox_amd", "ferrox_mi")
    assert os.access(exe, os.X_OK)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
