"""Pins the CPU oracle (oracle/osqp_oracle.c) before it is trusted.

Checks, each chosen so that a plausible slip in the oracle fails one of them:
  * solver level: the closed-form solutions, statuses and objective values the
    reference's own tests hold (tests/*/generate_problem.py, large_qp_data.c);
  * kernels: SpMV / transpose SpMV / symmetric-triu SpMV, norms, diag(A'DA)
    against scipy on random matrices (a dropped mirror term or a transposed
    operand changes the result);
  * structure: CSC->CSR and triu->full against scipy's conversions, bit-exact;
  * linear algebra: the LDL^T KKT solve against a dense numpy solve;
  * PCG restatement: converges to the same answer as the direct path.
No GPU is used.
"""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import OracleSolver, load_oracle
from osqp_amd import problems as pr
from osqp_amd._types import OSQP_NAN

from .conftest import GOLDEN

TESTS_TOL = 1e-4  # tests/osqp_tester.h:12-16 (double build)
FIXTURE = dict(rho=0.1, alpha=1.6, max_iter=2000, scaling=1, eps_abs=1e-5, eps_rel=1e-5)  # osqp_tester.h:71-82


def _solve(prob, linsys=1, **kw):
    s = OracleSolver()
    st = dict(FIXTURE)
    st.update(kw)
    rc = s.setup(prob["P"], prob["q"], prob["A"], prob["l"], prob["u"], linsys_solver=linsys, **st)
    return s, rc


@pytest.mark.parametrize("linsys", [1, 2])
def test_basic_qp_solution(linsys):
    p = pr.basic_qp()
    s, rc = _solve(p, linsys)
    assert rc == 0
    s.solve()
    assert s.info.status_val == 1
    assert np.max(np.abs(s.x - p["x"])) < TESTS_TOL
    assert np.max(np.abs(s.y - p["y"])) < TESTS_TOL
    assert abs(s.info.obj_val - p["obj"]) < TESTS_TOL


@pytest.mark.parametrize("linsys", [1, 2])
def test_basic_qp2_update_vectors(linsys):
    """tests/basic_qp2/test_basic_qp2.cpp: solve, then update q and u."""
    p = pr.basic_qp2()
    s, rc = _solve(p, linsys, eps_abs=1e-6, eps_rel=1e-6)
    assert rc == 0
    s.solve()
    assert s.info.status_val == 1
    assert np.max(np.abs(s.x - p["x"])) / np.max(np.abs(p["x"])) < TESTS_TOL
    assert abs(s.info.obj_val - p["obj"]) / abs(p["obj"]) < TESTS_TOL
    assert s.update_data_vec(q=p["q_new"], u=p["u_new"]) == 0
    s.solve()
    assert s.info.status_val == 1
    assert np.max(np.abs(s.x - p["x_new"])) / np.max(np.abs(p["x_new"])) < TESTS_TOL
    assert abs(s.info.obj_val - p["obj_new"]) / abs(p["obj_new"]) < TESTS_TOL


@pytest.mark.parametrize("name", ["basic_lp", "no_active_set", "unconstrained"])
def test_closed_forms(name):
    p = getattr(pr, name)()
    s, rc = _solve(p)
    assert rc == 0
    s.solve()
    assert s.info.status_val == 1
    assert np.max(np.abs(s.x - p["x"])) < 1e-3
    if "y" in p and len(p["y"]):
        assert np.max(np.abs(s.y - p["y"])) < 1e-3
    assert abs(s.info.obj_val - p["obj"]) < 1e-3


@pytest.mark.parametrize("case", [1, 2, 3, 4])
@pytest.mark.parametrize("linsys", [1, 2])
def test_primal_dual_infeasibility(case, linsys):
    p = pr.primal_dual_infeasibility(case)
    s, rc = _solve(p, linsys)
    assert rc == 0
    s.solve()
    assert s.info.status.decode() == p["status"]
    if case == 1:
        assert np.max(np.abs(s.x - p["x"])) < TESTS_TOL
        assert abs(s.info.obj_val - p["obj"]) < TESTS_TOL
    else:
        # certificates: NaN-filled solution, normalised certificate
        assert np.all(s.x == OSQP_NAN)
        cert = s.prim_inf_cert if "primal" in p["status"] else s.dual_inf_cert
        assert abs(np.max(np.abs(cert)) - 1.0) < 1e-12


def test_primal_infeasibility_random():
    p = pr.primal_infeasibility()
    s, rc = _solve(p)
    assert rc == 0
    s.solve()
    assert s.info.status.decode() == "primal infeasible"


def test_non_cvx():
    """tests/non_cvx/test_non_cvx.cpp:10-78: direct -> setup error, indirect -> status at solve."""
    p = pr.non_cvx()
    s, rc = _solve(p, 1)
    assert rc == 4 and s.info.status_val == 9
    # "Nonconvex: Solve" (test_non_cvx.cpp:50-78): sigma = 5 makes P + sigma I
    # definite, so setup succeeds and the divergence is caught at solve time
    for linsys in (1, 2):
        s, rc = _solve(p, linsys, adaptive_rho=0, sigma=5.0)
        assert rc == 0
        s.solve()
        assert s.info.status_val == 9
        assert s.info.obj_val == OSQP_NAN


@pytest.mark.parametrize("linsys", [1, 2])
def test_large_qp_fixture(linsys):
    """tests/large_qp/test_large_qp.cpp:31-37 on the committed data set."""
    p = pr.load_fixture(os.path.join(GOLDEN, "large_qp.json"))
    s, rc = _solve(p, linsys)
    assert rc == 0
    s.solve()
    assert s.info.status_val == 1
    assert abs(s.info.obj_val - p["obj"]) / abs(p["obj"]) < TESTS_TOL


def test_termination_fixed_iterations():
    """test_basic_qp.cpp:718-774: check_termination=0 runs exactly max_iter."""
    p = pr.basic_qp()
    s = OracleSolver()
    s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], max_iter=400, scaling=0, check_termination=0,
            warm_starting=0)
    s.solve()
    assert s.info.iter == 400
    assert s.info.status_val == 1
    assert np.max(np.abs(s.x - p["x"])) < TESTS_TOL


def test_update_rho_same_iterations():
    """test_basic_qp.cpp:776-887: update_rho(0.7) matches a fresh setup with rho=0.7."""
    p = pr.basic_qp()
    kw = dict(adaptive_rho=0, eps_abs=5e-5, eps_rel=5e-5, check_termination=1)
    a = OracleSolver()
    a.setup(p["P"], p["q"], p["A"], p["l"], p["u"], rho=0.7, **kw)
    a.solve()
    b = OracleSolver()
    b.setup(p["P"], p["q"], p["A"], p["l"], p["u"], rho=0.1, **kw)
    assert b.update_rho(0.7) == 0
    b.solve()
    assert a.info.iter == b.info.iter
    assert a.info.status_val == b.info.status_val == 1


def test_warm_start_from_solution():
    """test_basic_qp.cpp:945-993: warm start at the optimum converges in 1 iteration."""
    p = pr.basic_qp()
    s, _ = _solve(p, check_termination=1, adaptive_rho=0)
    s.solve()
    it = s.info.iter
    s.warm_start(x=np.zeros(2), y=np.zeros(4))
    s.solve()
    assert s.info.iter == it
    s.warm_start(x=p["x"])
    s.warm_start(y=p["y"])
    s.solve()
    assert s.info.iter == 1


UPDATE_SECTIONS = [
    # (name, P update, A update, held solution key): tests/update_matrices/
    # test_update_matrices.cpp:138-384, with index vectors, with all values
    # (null index, n = nnz) and with n = 0 meaning all
    ("P idx", ("idx",), None, "P_new"), ("P all", ("all",), None, "P_new"), ("P nnz0", ("zero",), None, "P_new"),
    ("A idx", None, ("idx",), "A_new"), ("A all", None, ("all",), "A_new"), ("A nnz0", None, ("zero",), "A_new"),
    ("PA idx", ("idx",), ("idx",), "PA_new"), ("PA all", ("all",), ("all",), "PA_new"),
    ("PA nnz0", ("zero",), ("zero",), "PA_new"),
]


def update_args(p, Pmode, Amode):
    """Keyword arguments of update_data_mat for one section."""
    kw = {}
    for key, mode, new in (("P", Pmode, p["Pu_new"]), ("A", Amode, p["A_new"])):
        if mode is None:
            continue
        kw[key + "x"] = new.data
        if mode[0] == "idx":
            kw[key + "x_idx"] = np.arange(new.nnz, dtype=np.int32)
        elif mode[0] == "zero":
            kw[key + "x_n0"] = True
    return kw


@pytest.mark.parametrize("linsys", [1, 2])
@pytest.mark.parametrize("section", range(len(UPDATE_SECTIONS)))
def test_update_data_mat_held_solutions(linsys, section):
    """oracle_update_data_mat (src/osqp_api.c:1309-1409): after the update the
    solve reaches the solution the reference's test holds (TESTS_TOL), and the
    same as a fresh setup on the new matrices."""
    name, Pm, Am, key = UPDATE_SECTIONS[section]
    p = pr.update_matrices_problem()
    st = dict(FIXTURE, max_iter=1000)
    s, rc = _solve(p, linsys, **st)
    assert rc == 0
    s.solve()
    assert s.info.status_val == 1
    assert np.max(np.abs(s.x - p["x"])) < TESTS_TOL
    assert np.max(np.abs(s.y - p["y"])) < TESTS_TOL
    assert s.update_data_mat(**update_args(p, Pm, Am)) == 0
    s.solve()
    assert s.info.status_val == 1, name
    assert np.max(np.abs(s.x - p["x_" + key])) < TESTS_TOL, name
    assert np.max(np.abs(s.y - p["y"])) < TESTS_TOL, name
    assert abs(s.info.obj_val - p["obj_" + key]) < TESTS_TOL, name
    fresh = dict(p, P=p["Pu_new"] if Pm else p["P"], A=p["A_new"] if Am else p["A"])
    f, _ = _solve(fresh, linsys, **st)
    f.solve()
    assert np.max(np.abs(s.x - f.x)) < TESTS_TOL
    assert abs(s.info.obj_val - f.info.obj_val) < TESTS_TOL


def test_update_data_mat_rejects_bad_sizes():
    """src/osqp_api.c:1335-1366: more values than nonzeros, or a partial
    update without an index vector, return 1 (P) or 2 (A)."""
    p = pr.update_matrices_problem()
    s, _ = _solve(p)
    nP, nA = p["P"].nnz, p["A"].nnz
    assert s.update_data_mat(Px=np.ones(nP + 1)) == 1
    assert s.update_data_mat(Px=np.ones(nP - 1)) == 1
    assert s.update_data_mat(Ax=np.ones(nA + 1)) == 2
    assert s.update_data_mat(Ax=np.ones(nA - 1)) == 2


def test_time_limit():
    p = pr.basic_qp()
    s, _ = _solve(p, rho=20, adaptive_rho=0)
    s.solve()
    assert s.info.status_val == 1
    s.update_settings(time_limit=1e-7, eps_abs=1e-12, eps_rel=1e-12, max_iter=2_000_000_000, check_termination=0)
    s.solve()
    assert s.info.status_val == 8


# ---- kernels ---------------------------------------------------------------------

def _rand(m, n, d, seed):
    rng = np.random.default_rng(seed)
    M = sp.random(m, n, density=d, format="csc", random_state=rng, data_rvs=rng.standard_normal)
    M.sort_indices()
    return M, rng


def _ptr(a):
    return a.ctypes.data


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_kernels_against_scipy(seed):
    L = load_oracle()
    A, rng = _rand(37, 23, 0.2, seed)
    Pf, _ = _rand(23, 23, 0.3, seed + 10)
    Pf = (Pf + Pf.T).tocsc()
    P = sp.triu(Pf, format="csc")
    P.sort_indices()
    x, y = rng.standard_normal(23), rng.standard_normal(37)
    Ap, Ai, Ax = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    Pp, Pi, Px = P.indptr.astype(np.int32), P.indices.astype(np.int32), P.data.copy()
    out = rng.standard_normal(37)
    ref = 2.0 * (A @ x) - 0.5 * out
    L.oracle_csc_Axpy(37, 23, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(x), _ptr(out), 2.0, -0.5)
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-13)
    out = np.zeros(23)
    L.oracle_csc_Atxpy(37, 23, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(y), _ptr(out), 1.0, 0.0)
    np.testing.assert_allclose(out, A.T @ y, rtol=1e-13, atol=1e-13)
    out = np.zeros(23)
    L.oracle_csc_Axpy_sym_triu(23, _ptr(Pp), _ptr(Pi), _ptr(Px), _ptr(x), _ptr(out), 1.0, 0.0)
    np.testing.assert_allclose(out, Pf @ x, rtol=1e-13, atol=1e-13)
    E = np.zeros(23)
    L.oracle_csc_col_norm_inf(37, 23, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(E))
    assert np.array_equal(E, np.abs(A).max(axis=0).toarray().ravel())
    E = np.zeros(37)
    L.oracle_csc_row_norm_inf(37, 23, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(E))
    assert np.array_equal(E, np.abs(A).max(axis=1).toarray().ravel())
    E = np.zeros(23)
    L.oracle_csc_row_norm_inf_sym_triu(23, _ptr(Pp), _ptr(Pi), _ptr(Px), _ptr(E))
    assert np.array_equal(E, np.abs(Pf).max(axis=1).toarray().ravel())
    # triu column norm (builtin semantics of col_norm_inf on P)
    E = np.zeros(23)
    L.oracle_csc_col_norm_inf(23, 23, _ptr(Pp), _ptr(Pi), _ptr(Px), _ptr(E))
    assert np.array_equal(E, np.abs(P).max(axis=0).toarray().ravel())
    D = rng.random(37) + 0.5
    d = np.zeros(23)
    L.oracle_csc_AtDA_extract_diag(37, 23, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(D), _ptr(d))
    np.testing.assert_allclose(d, (A.T @ sp.diags(D) @ A).diagonal(), rtol=1e-13)


@pytest.mark.parametrize("seed", [0, 3])
def test_structure_against_scipy(seed):
    L = load_oracle()
    A, _ = _rand(41, 29, 0.15, seed)
    Ap, Ai = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    rp, ci, pos = np.zeros(42, np.int32), np.zeros(A.nnz, np.int32), np.zeros(A.nnz, np.int32)
    L.oracle_csc_to_csr(41, 29, _ptr(Ap), _ptr(Ai), _ptr(rp), _ptr(ci), _ptr(pos))
    R = A.tocsr()
    R.sort_indices()
    assert np.array_equal(rp, R.indptr) and np.array_equal(ci, R.indices)
    vals = np.zeros(A.nnz)
    vals[pos] = A.data
    assert np.array_equal(vals, R.data)
    Pf, _ = _rand(31, 31, 0.2, seed + 5)
    Pf = (Pf + Pf.T + sp.eye(31) * 0.0).tocsc()
    P = sp.triu(Pf, format="csc")
    P.sort_indices()
    Pp, Pi = P.indptr.astype(np.int32), P.indices.astype(np.int32)
    F = (P + sp.triu(P, 1).T).tocsr()
    F.sort_indices()
    rp, ci = np.zeros(32, np.int32), np.zeros(F.nnz, np.int32)
    fpos, mpos = np.zeros(P.nnz, np.int32), np.zeros(P.nnz, np.int32)
    L.oracle_triu_to_full_csr(31, _ptr(Pp), _ptr(Pi), _ptr(rp), _ptr(ci), _ptr(fpos), _ptr(mpos))
    assert np.array_equal(rp, F.indptr) and np.array_equal(ci, F.indices)
    vals = np.zeros(F.nnz)
    vals[fpos] = P.data
    vals[mpos[mpos >= 0]] = P.data[mpos >= 0]
    assert np.array_equal(vals, F.data)


@pytest.mark.parametrize("seed", [0, 1])
def test_ldl_kkt_solve_against_dense(seed):
    """(P + sigma I + A' R A) x = b solved through the quasi-definite KKT."""
    L = load_oracle()
    rng = np.random.default_rng(seed)
    n, m = 30, 45
    M = sp.random(n, n, density=0.2, random_state=rng)
    Pf = (M @ M.T + 0.1 * sp.eye(n)).tocsc()
    P = sp.triu(Pf, format="csc")
    A = sp.random(m, n, density=0.2, format="csc", random_state=rng)
    rho = rng.random(m) + 0.1
    sigma = 1e-6
    b = np.hstack([rng.standard_normal(n), np.zeros(m)])
    K = Pf.toarray() + sigma * np.eye(n) + A.T.toarray() @ np.diag(rho) @ A.toarray()
    x_ref = np.linalg.solve(K, b[:n])
    Pp, Pi, Px = P.indptr.astype(np.int32), P.indices.astype(np.int32), P.data.copy()
    Ap, Ai, Ax = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    rc = L.oracle_kkt_solve(n, m, _ptr(Pp), _ptr(Pi), _ptr(Px), _ptr(Ap), _ptr(Ai), _ptr(Ax), sigma,
                            _ptr(rho), _ptr(b))
    assert rc == 0
    np.testing.assert_allclose(b[:n], x_ref, rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("seed", [0, 1])
def test_pcg_matches_direct(seed):
    p = pr.random_qp(60, 300, 0.1, 0.1, seed)
    d, _ = _solve(p, 1, eps_abs=1e-5, eps_rel=1e-5)
    d.solve()
    i, _ = _solve(p, 2, eps_abs=1e-5, eps_rel=1e-5)
    i.solve()
    assert d.info.status_val == i.info.status_val == 1
    assert abs(d.info.obj_val - i.info.obj_val) / abs(d.info.obj_val) < 1e-4
    assert i.total_pcg_iters() > 0
