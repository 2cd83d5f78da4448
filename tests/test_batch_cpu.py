"""Host-side pieces of the batching form (no GPU): the spectral basis that
diagonalises the periodic lateral operator, and the batching workload
generator."""
import numpy as np
import pytest

import ferrox_amd as fx
from ferrox_amd import configs
from ferrox_amd.batch import basis


@pytest.mark.parametrize("n", [1, 2, 3, 5, 32, 48, 64, 96, 128])
def test_basis_is_orthonormal_and_diagonalises_the_periodic_second_difference(n):
    Q, mu = basis(n)
    L = 2 * np.eye(n) - np.roll(np.eye(n), 1, 0) - np.roll(np.eye(n), -1, 0)
    if n == 1:
        L = np.zeros((1, 1))
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-13
    assert np.abs(Q.T @ L @ Q - np.diag(mu)).max() < 1e-12 * max(1, n)
    # the spectrum is the textbook one: 2 - 2 cos(2 pi f / n), f = 0..n-1
    expect = np.sort(2 - 2 * np.cos(2 * np.pi * np.arange(n) / n))
    assert np.allclose(np.sort(mu), expect, atol=1e-14)


def test_batching_workload_matches_its_specification():
    texts = configs.batch_devices(n=48, nsteps=5)
    assert len(texts) == 48
    seen = set()
    for t in texts:
        inp = fx.from_text(t)
        nx, ny, nz = inp.n_cell
        assert nx in (32, 48, 64, 96, 128) and ny in (32, 48, 64, 96, 128)
        dz = (inp.prob_hi[2] - inp.prob_lo[2]) / nz
        assert abs(dz - 0.25e-9) < 1e-18
        nfe = nz - 16
        assert 12 <= nfe <= 40
        assert abs(inp.DE_hi[2] - 4e-9) < 1e-18
        assert -1.0 <= inp.Phi_Bc_hi <= 1.0
        seen.add(inp.random_seed)
    assert len(seen) == 48
