"""The two configurations too large for the oracle inside a test, at full size,
against the oracle's FP64 PCG answers held in tests/golden (written once by
tests/golden/make_full_size_fixtures.py: the same generator and seed, the
oracle's restatement of the reference's PCG path, linsys_solver = 2):

  * large random QP: n = 50000, m = 500000, A of density 2e-3 (5.0e7
    nonzeros), P = M M' + 1e-2 I, eps 1e-4;
  * lasso: n = 20000, m_d = 200000, D of density 1e-3 (QP of 240000
    variables and constraints, 4.3e6 nonzeros in A), eps 1e-4.

Both HIP engines must reach the oracle's status; the plain FP64 engine runs
the same algorithm in the same arithmetic, so its ADMM iteration count and rho
updates equal the oracle's and its point agrees to rounding (objective 1e-9,
x and y 1e-6 relative); the FP32-PCG engine (FP32 CG values, FP64 defect
correction and residuals) agrees to the stated FP32 bounds (objective 1e-6,
x and y 1e-3 relative).  The generator's data checksums are compared first,
so a changed generator fails loudly instead of comparing different problems.
"""
import os

import numpy as np
import pytest

from osqp_amd import OSQP
from osqp_amd import problems as pr

from .conftest import GOLDEN
from .oracle_refs import rel, vec_err

pytestmark = pytest.mark.gpu

CONFIGS = {
    "large_qp": lambda: pr.random_qp(50000, 500000, 1e-4, 2e-3, seed=0),
    "lasso": lambda: pr.lasso(20000, 200000, 1e-3, seed=0),
}
BOUNDS = {64: dict(obj=1e-9, xy=1e-6), 32: dict(obj=1e-6, xy=1e-3)}


def checksums(p):
    return np.array([p["P"].nnz, p["A"].nnz, np.sum(p["P"].data), np.sum(p["A"].data), np.sum(p["q"]),
                     np.sum(np.clip(p["l"], -1e30, 1e30)), np.sum(np.clip(p["u"], -1e30, 1e30))])


_cache = {}


def problem(name):
    if name not in _cache:
        z = np.load(os.path.join(GOLDEN, f"{name}_oracle.npz"))
        p = CONFIGS[name]()
        np.testing.assert_allclose(checksums(p), z["checksums"], rtol=1e-12, atol=0)
        _cache.clear()
        _cache[name] = (p, z)
    return _cache[name]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", ["large_qp", "lasso"])
def test_full_size_vs_oracle(name, precision):
    p, z = problem(name)
    info = z["info"]
    o_status, o_iter, o_rho_updates, o_obj = int(info[0]), int(info[1]), int(info[2]), float(info[3])
    print(f"{name} p{precision}: oracle status {o_status} iter {o_iter} rho_updates {o_rho_updates} obj {o_obj:.12e}",
          flush=True)
    s = OSQP(precision=precision)
    assert s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], eps_abs=1e-4, eps_rel=1e-4, verbose=0) == 0
    r = s.solve()
    s.close()
    e_obj, e_x, e_y = rel(r.info.obj_val, o_obj), vec_err(r.x, z["x"]), vec_err(r.y, z["y"])
    print(f"{name} p{precision}: status {r.info.status_val} iter {r.info.iter} rho_updates {r.info.rho_updates} "
          f"obj {r.info.obj_val:.12e}: obj {e_obj:.1e} x {e_x:.1e} y {e_y:.1e}")
    assert r.info.status_val == o_status == 1
    if precision == 64:
        assert r.info.iter == o_iter and r.info.rho_updates == o_rho_updates
    b = BOUNDS[precision]
    assert e_obj <= b["obj"]
    assert e_x <= b["xy"] and e_y <= b["xy"]
