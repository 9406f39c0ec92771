"""The PCG's operators in each of their kernel forms, against scipy.

K = P + sigma I + A' diag(rho) A (the reduced KKT operator, cuda_pcg.cu:50-106)
and the product A x are applied to a random vector through the solver's own
kernels (osqp_amd_apply_operator) on three operators that take the three
forms:
  * column panels (a 30000 x 9000 A with 30 entries per row: K spans 39000
    columns, 5 FP32 / 10 FP64 panels);
  * hybrids (the lasso QP with n = 1000, m_d = 10000: panels over D's rows and
    columns, a remainder CSR for the identity blocks);
  * row-block CSR (a 60000 x 30000 A with 1.5 entries per row: neither
    panels nor a hybrid pays).
Scaling is off, so the solver's P, A are the inputs; rho per constraint type
as set_rho_vec gives it (src/auxil.c:76-122).  Bounds, elementwise, with
|K||x| the product of the absolute values: FP64 vectors and values 16 eps;
FP64 vectors with the FP32-rounded values of the FP32 engine's defect
correction 2^-22; the FP32 CG operator 2^-16 (FP32 products summed in FP32).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from osqp_amd import OSQP
from osqp_amd import problems as pr

pytestmark = pytest.mark.gpu

RHO, SIGMA = 0.1, 1e-6
OPS = {
    "panels": lambda: pr.random_qp(9000, 30000, 2.0 / 9000, 30.0 / 9000, seed=7),
    "hybrid": lambda: pr.lasso(1000, 10000, 0.01, seed=1),
    "row_block": lambda: pr.random_qp(30000, 60000, 1e-5, 5e-5, seed=3),
}


def rho_vec(l, u):
    l, u = np.clip(l, -1e30, 1e30), np.clip(u, -1e30, 1e30)
    infval = 1e30 * 1e-4
    r = np.full(len(l), RHO)
    r[u - l < 1e-4] = 1e3 * RHO
    r[(l < -infval) & (u > infval)] = 1e-6
    return r


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", sorted(OPS))
def test_operator_forms_vs_scipy(name, precision):
    p = OPS[name]()
    n, m = p["P"].shape[0], p["A"].shape[0]
    s = OSQP(precision=precision)
    assert s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], scaling=0, rho=RHO, sigma=SIGMA, adaptive_rho=0) == 0
    P = sp.csc_matrix(p["P"])
    Pf = P + sp.triu(P, 1).T
    A = sp.csc_matrix(p["A"])
    r = rho_vec(p["l"], p["u"])
    x = np.random.default_rng(11).standard_normal(n)
    K = Pf + SIGMA * sp.eye(n) + A.T @ sp.diags(r) @ A
    Kx = K @ x
    Kabs = abs(Pf) @ abs(x) + SIGMA * abs(x) + abs(A).T @ (r * (abs(A) @ abs(x)))
    Ax, Aabs = A @ x, abs(A) @ abs(x)
    eps64, eps_dc, eps32 = 16 * np.finfo(float).eps, 2.0 ** -22, 2.0 ** -16
    checks = [(0, Kx, Kabs, eps64 if precision == 64 else eps32),
              (1, Kx, Kabs, eps64 if precision == 64 else eps_dc),
              (2, Ax, Aabs, eps64 if precision == 64 else eps_dc)]
    for which, ref, mag, eps in checks:
        y = s.apply_operator(which, x)
        err = np.abs(y - ref) / np.maximum(mag, 1e-300)
        print(f"{name} p{precision} which {which}: max |y - ref| / (|M||x|) = {err.max():.2e} (bound {eps:.1e})")
        assert np.all(np.abs(y - ref) <= eps * mag + 1e-300)
    s.close()
