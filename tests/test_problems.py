"""Workload generators: the MPC stream's per-horizon template gives the same
matrices, entry for entry, as building each problem with scipy."""
import numpy as np
import pytest

from osqp_amd import problems as pr


@pytest.mark.parametrize("N", [10, 11, 37, 50])
@pytest.mark.parametrize("seed", [0, 3, 1001])
def test_mpc_template_equals_scipy_build(N, seed):
    a = pr.mpc(N, seed=seed)
    rng = np.random.default_rng(seed)
    nx, nu = 12, 4
    Ad = rng.standard_normal((nx, nx))
    Ad *= 0.95 / max(abs(np.linalg.eigvals(Ad)))
    Bd = rng.standard_normal((nx, nu))
    x0 = rng.uniform(-1, 1, nx)
    b = pr.mpc_build(N, Ad, Bd, x0)
    for k in ("P", "A"):
        assert a[k].shape == b[k].shape
        np.testing.assert_array_equal(a[k].indptr, b[k].indptr)
        np.testing.assert_array_equal(a[k].indices, b[k].indices)
        np.testing.assert_array_equal(a[k].data, b[k].data)
    for k in ("q", "l", "u"):
        np.testing.assert_array_equal(a[k], b[k])
    # shapes of the MPC configuration: n = 16N + 12, m = 28N + 24, nnz(A) = 220N + 24
    assert a["A"].shape == (28 * N + 24, 16 * N + 12) and a["A"].nnz == 220 * N + 24
