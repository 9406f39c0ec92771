"""Integer and scaling work of the HIP backend against the oracle, exactly.

  * structure (csc_utils.c:247-286 triplet_to_csr; cuda_csr.cu:489-577
    triu->full and transpose): the backend's CSR of A, its A' and the full
    symmetric CSR of P, with the maps from the input CSC entries, equal the
    oracle's conversions (oracle_csc_to_csr / oracle_triu_to_full_csr, pinned
    against scipy in test_oracle.py) array for array, and the stored values
    sit where the maps say;
  * Ruiz scaling (src/scaling.c:48-164): D and E are bit-exact with the
    oracle's whenever the cost scaling c is (its mean column norm is a sum
    whose order differs: scaling = 1 fixes D and E before c is used, and a
    q-dominated problem takes c from ||q||_inf, a max); otherwise c agrees
    within the norm_1 bound and D, E to a few ulps per Ruiz pass.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import OracleSolver, load_oracle
from osqp_amd import OSQP
from osqp_amd import problems as pr
from osqp_amd.solver import matrix_export

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps


def oracle_csr(A):
    O = load_oracle()
    m, n = A.shape
    p, i = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    rp, ci, pos = np.zeros(m + 1, np.int32), np.zeros(max(A.nnz, 1), np.int32), np.zeros(max(A.nnz, 1), np.int32)
    O.oracle_csc_to_csr(m, n, p.ctypes.data, i.ctypes.data, rp.ctypes.data, ci.ctypes.data, pos.ctypes.data)
    return rp, ci[:A.nnz], pos[:A.nnz]


def oracle_full(P):
    O = load_oracle()
    n = P.shape[0]
    p, i = P.indptr.astype(np.int32), P.indices.astype(np.int32)
    full = P.nnz + sum(1 for j in range(n) for k in range(P.indptr[j], P.indptr[j + 1]) if P.indices[k] != j)
    rp, ci = np.zeros(n + 1, np.int32), np.zeros(max(full, 1), np.int32)
    pos, mir = np.zeros(max(P.nnz, 1), np.int32), np.zeros(max(P.nnz, 1), np.int32)
    O.oracle_triu_to_full_csr(n, p.ctypes.data, i.ctypes.data, rp.ctypes.data, ci.ctypes.data, pos.ctypes.data,
                              mir.ctypes.data)
    return rp, ci[:full], pos[:P.nnz], mir[:P.nnz]


def structure_cases():
    rng = np.random.default_rng(5)
    out = []
    # empty rows and columns, a dense row, a dense column
    A = sp.random(300, 120, density=0.05, format="lil", random_state=rng)
    A[7, :] = 1.0
    A[:, 11] = 2.0
    A[20, :] = 0.0
    A[:, 50] = 0.0
    out.append(("ragged", sp.csc_matrix(A)))
    out.append(("config_shape", pr.random_qp(2000, 20000, 1e-3, 2e-3, seed=1)["A"]))
    out.append(("one_row", sp.csc_matrix(rng.standard_normal((1, 40)))))
    out.append(("one_col", sp.csc_matrix(rng.standard_normal((40, 1)))))
    return out


@pytest.mark.parametrize("case", range(4))
def test_csr_and_transpose_bitexact(case):
    name, A = structure_cases()[case]
    A = sp.csc_matrix(A)
    A.sort_indices()
    m, n = A.shape
    P = sp.eye(n, format="csc")
    s = OSQP(precision=64)
    assert s.setup(P, np.zeros(n), A, -np.ones(m), np.ones(m), scaling=0) == 0
    st = s.structures()["A"]
    rp, ci, pos = oracle_csr(A)
    np.testing.assert_array_equal(st["rp"], rp)
    np.testing.assert_array_equal(st["ci"], ci)
    np.testing.assert_array_equal(st["csc2csr"], pos)
    # values where the map says, and A' stored as the input CSC
    np.testing.assert_array_equal(st["val"][pos], A.data)
    np.testing.assert_array_equal(st["trp"], A.indptr)
    np.testing.assert_array_equal(st["tci"], A.indices)
    np.testing.assert_array_equal(st["tval"], A.data)
    # and the CSR is scipy's
    R = A.tocsr()
    R.sort_indices()
    np.testing.assert_array_equal(st["rp"], R.indptr)
    np.testing.assert_array_equal(st["ci"], R.indices)
    np.testing.assert_array_equal(st["val"], R.data)
    s.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_triu_to_full_bitexact(seed):
    """P given by its upper triangle, with some diagonal entries missing."""
    rng = np.random.default_rng(seed)
    n = 200
    M = sp.random(n, n, density=0.05, random_state=rng, data_rvs=rng.standard_normal)
    Pf = (M @ M.T).tocsc()
    P = sp.triu(Pf, format="csc")
    keep = np.ones(P.nnz, bool)
    for j in range(0, n, 7):  # drop every seventh diagonal entry
        for k in range(P.indptr[j], P.indptr[j + 1]):
            if P.indices[k] == j:
                keep[k] = False
    P = sp.csc_matrix(sp.coo_matrix((P.tocoo().data[keep], (P.tocoo().row[keep], P.tocoo().col[keep])), shape=(n, n)))
    P.sort_indices()
    A = sp.random(50, n, density=0.1, format="csc", random_state=rng)
    s = OSQP(precision=64)
    assert s.setup(P, np.zeros(n), A, -np.ones(50), np.ones(50), scaling=0) == 0
    st = s.structures()["P"]
    rp, ci, pos, mir = oracle_full(P)
    np.testing.assert_array_equal(st["rp"], rp)
    np.testing.assert_array_equal(st["ci"], ci)
    np.testing.assert_array_equal(st["csc2csr"], pos)
    np.testing.assert_array_equal(st["mirror"], mir)
    np.testing.assert_array_equal(st["val"][pos], P.data)
    off = mir >= 0
    np.testing.assert_array_equal(st["val"][mir[off]], P.data[off])
    # diagonal slots: (j, j) where stored, -1 where P has no diagonal entry
    diag_ref = np.full(n, -1, np.int32)
    for k in range(P.nnz):
        if mir[k] < 0:
            j = int(np.searchsorted(P.indptr, k, side="right") - 1)
            diag_ref[j] = pos[k]
    np.testing.assert_array_equal(st["diag"], diag_ref)
    # the full CSR is scipy's full symmetric matrix
    F = (P + sp.triu(P, 1).T).tocsr()
    F.sort_indices()
    np.testing.assert_array_equal(st["rp"], F.indptr)
    np.testing.assert_array_equal(st["ci"], F.indices)
    np.testing.assert_array_equal(st["val"], F.data)
    s.close()


def _scaling_pair(p, scaling):
    st = dict(scaling=scaling, eps_abs=1e-4, eps_rel=1e-4)
    o = OracleSolver()
    assert o.setup(p["P"], p["q"], p["A"], p["l"], p["u"], linsys_solver=2, **st) == 0
    s = OSQP(precision=64)
    assert s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **st) == 0
    return o, s


@pytest.mark.parametrize("name", ["random_qp", "mpc", "portfolio", "lasso"])
def test_ruiz_one_pass_bitexact(name):
    """scaling = 1: D and E come from the raw data's column and row norms,
    square roots and reciprocals only, so they are bit-exact; c, the scaled
    q, l, u and the scaled matrices follow (c within the norm_1 bound)."""
    p = dict(random_qp=lambda: pr.random_qp(80, 400, 0.1, 0.1, seed=3), mpc=lambda: pr.mpc(20, seed=4),
             portfolio=lambda: pr.portfolio(300, seed=5), lasso=lambda: pr.lasso(40, 300, 0.1, seed=6))[name]()
    o, s = _scaling_pair(p, 1)
    Do, Eo, co, qo, lo, uo = o.scaling()
    Dg, Eg, cg, qg, lg, ug = s.scaling()
    np.testing.assert_array_equal(Dg, Do)
    np.testing.assert_array_equal(Eg, Eo)
    n = len(Do)
    assert abs(cg - co) <= 2 * n * EPS * abs(co), (cg, co)
    np.testing.assert_array_equal(lg, lo)
    np.testing.assert_array_equal(ug, uo)
    if cg == co:
        np.testing.assert_array_equal(qg, qo)
    s.close()


def q_dominated(seed):
    """A QP whose ||q||_inf exceeds the mean column norm of P in every Ruiz
    pass, so c = 1/||q||_inf exactly (a max, no sum)."""
    p = pr.random_qp(60, 300, 0.1, 0.1, seed=seed)
    p["q"] = 1e3 * p["q"]
    return p


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_ruiz_ten_passes_bitexact_when_c_is(seed):
    p = q_dominated(seed)
    o, s = _scaling_pair(p, 10)
    Do, Eo, co, qo, lo, uo = o.scaling()
    Dg, Eg, cg, qg, lg, ug = s.scaling()
    assert cg == co
    np.testing.assert_array_equal(Dg, Do)
    np.testing.assert_array_equal(Eg, Eo)
    np.testing.assert_array_equal(qg, qo)
    np.testing.assert_array_equal(lg, lo)
    np.testing.assert_array_equal(ug, uo)
    Pxo, Axo = o.scaled_matrices()
    st = s.structures()
    np.testing.assert_array_equal(st["P"]["val"][st["P"]["csc2csr"]], Pxo[:len(st["P"]["csc2csr"])])
    np.testing.assert_array_equal(st["A"]["tval"], Axo[:len(st["A"]["tval"])])
    s.close()


@pytest.mark.parametrize("name", ["random_qp", "portfolio"])
def test_ruiz_ten_passes_general(name):
    """Ten passes where c comes from P's mean column norm (a sum): c within
    the norm_1 bound per pass, D and E within a few ulps per pass."""
    p = dict(random_qp=lambda: pr.random_qp(80, 400, 0.1, 0.1, seed=9),
             portfolio=lambda: pr.portfolio(400, seed=2))[name]()
    o, s = _scaling_pair(p, 10)
    Do, Eo, co, *_ = o.scaling()
    Dg, Eg, cg, *_ = s.scaling()
    n = len(Do)
    assert abs(cg - co) <= 10 * 2 * n * EPS * abs(co)
    np.testing.assert_allclose(Dg, Do, rtol=200 * EPS, atol=0)
    np.testing.assert_allclose(Eg, Eo, rtol=200 * EPS, atol=0)
    s.close()
