"""fx_poisson_solve_noalloc (the direct solve on device arrays, workspace in
a static device array) against fx_poisson_solve on the same context, through
the C ABI: bitwise equal, on two calls in a row (the second on a smaller
problem, so it reuses the front of the workspace the first one filled), and
the size limit and the direct-path requirement rejected with FX_EINVAL.
"""
import ctypes as C

import numpy as np
import pytest
# torch first: its wheel carries its own HIP runtime under the system
# runtime's soname (libamdhip64.so.7), and whichever is loaded first serves
# the whole process; torch finds no GPU on the system runtime, the library
# runs on either
import torch  # noqa: F401

import ferrox_amd as fx
from ferrox_amd import configs

pytestmark = pytest.mark.gpu

FX_EINVAL = 1


def solve_both(text, seed):
    inp = fx.from_text(text, mg_tol_rel=1e-13)
    g = fx.FerroX(inp)
    lib = g._lib
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(g.shape) * 1e6
    ref, st = g.poisson_solve(b)
    assert st["vcycles"] == 0                               # the direct path
    rhs_d = torch.tensor(b, dtype=torch.float64, device="cuda")
    phi_d = torch.full(g.shape, float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()                                # torch's stream is not the context's
    rc = lib.fx_poisson_solve_noalloc(g._h, C.c_void_p(rhs_d.data_ptr()), C.c_void_p(phi_d.data_ptr()))
    assert rc == 0, lib.fx_last_error(g._h)
    g.sync()
    out = phi_d.cpu().numpy()
    g.close()
    return out, ref


def test_noalloc_matches_poisson_solve_two_calls():
    big = configs.mfim(nx=128, ny=128, nz=64, lx=64e-9, ly=64e-9, lz=16e-9, de_top=6e-9, fe_top=16e-9, bias=0.4)
    small = configs.mfim(nx=64, ny=32, nz=40, lx=32e-9, ly=16e-9, lz=10e-9, de_top=4e-9, fe_top=10e-9, bias=-0.3)
    for name, text, seed in (("128x128x64", big, 1), ("64x32x40", small, 2)):
        out, ref = solve_both(text, seed)
        print(name, "max |noalloc - poisson_solve|", np.abs(out - ref).max(), flush=True)
        assert np.array_equal(out, ref), name


def test_noalloc_rejects_what_it_cannot_do():
    buf = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = C.c_void_p(buf.data_ptr())
    # off the direct path: a lateral extent that is not a power of two
    g = fx.FerroX(fx.from_text(configs.mfim(nx=48, ny=32, nz=36, lx=24e-9, ly=16e-9, nsteps=1)))
    assert g._lib.fx_poisson_solve_noalloc(g._h, p, p) == FX_EINVAL
    print(g._lib.fx_last_error(g._h), flush=True)
    g.close()
    # beyond the 64 MiB workspace (N doubles, at most 256 x 256 x 128 cells):
    # 256 x 256 x 144 cells need 75.5 MB
    g = fx.FerroX(fx.from_text(configs.mfim(nx=256, ny=256, nz=144, lx=128e-9, ly=128e-9, lz=14.4e-9, de_top=4e-9,
                                            fe_top=12e-9, nsteps=1)))
    assert g._lib.fx_poisson_solve_noalloc(g._h, p, p) == FX_EINVAL
    msg = g._lib.fx_last_error(g._h)
    print(msg, flush=True)
    assert b"static array" in msg
    g.close()
