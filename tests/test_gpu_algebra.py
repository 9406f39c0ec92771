"""Algebra-boundary parity: every algebra/hip symbol against the CPU oracle.

Mirrors the reference's lin_alg tests (tests/lin_alg/test_mat_vec.cpp,
test_matrix.cpp, testcases/test_vector_math_ops.cpp), which test "at the API
between OSQP and a linear algebra module" (tests/CMakeLists.txt:81-93).

Bar: elementwise ops, max-type reductions, comparisons, classification and all
structure work are bit-exact (compared with ==); sums and SpMVs are compared
with a bound of 4 * nnz_row * eps * (|A||x|) (order of summation differs).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import load_oracle
from osqp_amd.algebra import IVec, Mat, Vec

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def rng_vecs(seed, n, k=3):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(n) for _ in range(k)], rng


@pytest.mark.parametrize("n", [0, 1, 7, 1000, 100_003])
def test_elementwise_bitexact(algebra, n):
    (a, b, c), rng = rng_vecs(n, n)
    L = algebra.lib()
    va, vb, vc, vx = Vec(a), Vec(b), Vec(c), Vec(n=n)
    L.OSQPVectorf_add_scaled(vx.h, 0.3, va.h, -1.7, vb.h)
    assert np.array_equal(vx.get(), 0.3 * a + -1.7 * b)
    L.OSQPVectorf_add_scaled3(vx.h, 0.3, va.h, -1.7, vb.h, 2.5, vc.h)
    assert np.array_equal(vx.get(), 0.3 * a + -1.7 * b + 2.5 * c)
    # increment form x += sb*b + sc*c  (vector.c:439-466)
    x0 = vx.get()
    L.OSQPVectorf_add_scaled3(vx.h, 1.0, vx.h, -1.7, vb.h, 2.5, vc.h)
    assert np.array_equal(vx.get(), x0 + (-1.7 * b + 2.5 * c))
    L.OSQPVectorf_ew_prod(vx.h, va.h, vb.h)
    assert np.array_equal(vx.get(), a * b)
    L.OSQPVectorf_plus(vx.h, va.h, vb.h)
    assert np.array_equal(vx.get(), a + b)
    L.OSQPVectorf_minus(vx.h, va.h, vb.h)
    assert np.array_equal(vx.get(), a - b)
    L.OSQPVectorf_ew_max_vec(vx.h, va.h, vb.h)
    assert np.array_equal(vx.get(), np.where(a > b, a, b))
    L.OSQPVectorf_ew_min_vec(vx.h, va.h, vb.h)
    assert np.array_equal(vx.get(), np.where(a < b, a, b))
    vp = Vec(np.abs(a) + 0.1)
    L.OSQPVectorf_ew_sqrt(vp.h)
    assert np.array_equal(vp.get(), np.sqrt(np.abs(a) + 0.1))
    L.OSQPVectorf_ew_reciprocal(vx.h, vp.h)
    assert np.array_equal(vx.get(), 1.0 / np.sqrt(np.abs(a) + 0.1))
    lo, hi = np.minimum(b, c), np.maximum(b, c)
    vl, vu = Vec(lo), Vec(hi)
    L.OSQPVectorf_ew_bound_vec(vx.h, va.h, vl.h, vu.h)
    assert np.array_equal(vx.get(), np.where(np.where(a > lo, a, lo) < hi, np.where(a > lo, a, lo), hi))
    L.OSQPVectorf_copy(vx.h, va.h)
    L.OSQPVectorf_mult_scalar(vx.h, -3.25)
    assert np.array_equal(vx.get(), a * -3.25)
    L.OSQPVectorf_set_scalar_if_lt(vx.h, va.h, 0.1, 1.0)
    assert np.array_equal(vx.get(), np.where(a < 0.1, 1.0, a))
    L.OSQPVectorf_set_scalar_if_gt(vx.h, va.h, 0.1, 7.0)
    assert np.array_equal(vx.get(), np.where(a > 0.1, 7.0, a))
    small = a * 1e-16
    vs = Vec(small)
    L.OSQPVectorf_round_to_zero(vs.h, 1e-15)
    assert np.array_equal(vs.get(), np.where(np.abs(small) < 1e-15, 0.0, small))
    for v in (va, vb, vc, vx, vp, vl, vu, vs):
        v.free()


@pytest.mark.parametrize("n", [1, 1000, 65_537])
def test_reductions(algebra, n):
    (a, b, s), rng = rng_vecs(n + 1, n)
    L = algebra.lib()
    va, vb, vs = Vec(a), Vec(b), Vec(s)
    # max-type reductions are exact in any order
    assert L.OSQPVectorf_norm_inf(va.h) == np.max(np.abs(a))
    assert L.OSQPVectorf_scaled_norm_inf(vs.h, va.h) == np.max(np.abs(s * a))
    assert L.OSQPVectorf_norm_inf_diff(va.h, vb.h) == np.max(np.abs(a - b))
    tol = 4 * n * EPS
    assert abs(L.OSQPVectorf_norm_1(va.h) - np.sum(np.abs(a))) <= tol * np.sum(np.abs(a))
    assert abs(L.OSQPVectorf_dot_prod(va.h, vb.h) - a @ b) <= tol * (np.abs(a) @ np.abs(b))
    pos = a @ np.maximum(b, 0)
    neg = a @ np.minimum(b, 0)
    assert abs(L.OSQPVectorf_dot_prod_signed(va.h, vb.h, 1) - pos) <= tol * (np.abs(a) @ np.abs(b))
    assert abs(L.OSQPVectorf_dot_prod_signed(va.h, vb.h, -1) - neg) <= tol * (np.abs(a) @ np.abs(b))
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    vl, vu = Vec(lo), Vec(hi)
    assert L.OSQPVectorf_all_leq(vl.h, vu.h) == 1
    assert L.OSQPVectorf_all_leq(vu.h, vl.h) == (1 if np.all(hi <= lo) else 0)
    assert L.OSQPVectorf_is_eq(va.h, va.h, 0.0) == 1
    assert L.OSQPVectorf_is_eq(va.h, vb.h, 1e-9) == 0
    for v in (va, vb, vs, vl, vu):
        v.free()


@pytest.mark.parametrize("n", [0, 1, 1000, 300_001])
def test_queued_reductions_match(algebra, n):
    """Queued reductions (run together in one launch when waited for) give
    the bits of the synchronous calls, including past the 16-entry queue
    (which runs the queue first) and for empty vectors."""
    import ctypes as C
    (a, b, s), _ = rng_vecs(n + 7, n)
    L = algebra.lib()
    va, vb, vs = Vec(a), Vec(b), Vec(s)
    calls = [("norm", (va,)), ("scaled", (vs, va)), ("dot", (va, vb)), ("signed", (va, vb, 1)),
             ("signed", (va, vb, -1)), ("norm", (vb,)), ("scaled", (va, vb)), ("dot", (vb, vs))] * 3
    outs = (C.c_double * len(calls))()
    for k, (kind, args) in enumerate(calls):
        dst = C.cast(C.addressof(outs) + 8 * k, C.c_void_p)
        if kind == "norm":
            L.osqp_amd_red_norm_inf(args[0].h, dst)
        elif kind == "scaled":
            L.osqp_amd_red_scaled_norm_inf(args[0].h, args[1].h, dst)
        elif kind == "dot":
            L.osqp_amd_red_dot_prod(args[0].h, args[1].h, dst)
        else:
            L.osqp_amd_red_dot_prod_signed(args[0].h, args[1].h, args[2], dst)
    L.osqp_amd_red_wait()
    for k, (kind, args) in enumerate(calls):
        if kind == "norm":
            ref = L.OSQPVectorf_norm_inf(args[0].h)
        elif kind == "scaled":
            ref = L.OSQPVectorf_scaled_norm_inf(args[0].h, args[1].h)
        elif kind == "dot":
            ref = L.OSQPVectorf_dot_prod(args[0].h, args[1].h)
        else:
            ref = L.OSQPVectorf_dot_prod_signed(args[0].h, args[1].h, args[2])
        assert outs[k] == ref, (k, kind, outs[k], ref)
    for v in (va, vb, vs):
        v.free()


def test_reccone_and_bounds_type(algebra):
    """vector.c:676-727 and 885-921 with infinite, equal and loose bounds."""
    rng = np.random.default_rng(5)
    n = 5000
    INF = 1e30
    l = rng.standard_normal(n)
    u = l + np.abs(rng.standard_normal(n))
    kind = rng.integers(0, 5, n)
    l[kind == 0] = -INF
    u[kind == 1] = INF
    l[kind == 2], u[kind == 2] = -INF, INF
    u[kind == 3] = l[kind == 3] + 1e-5  # "equality" (u - l < 1e-4)
    y = rng.standard_normal(n)
    infval = 1e30 * 1e-4
    L = algebra.lib()
    vy, vl, vu = Vec(y), Vec(l), Vec(u)
    L.OSQPVectorf_project_polar_reccone(vy.h, vl.h, vu.h, infval)
    ref = y.copy()
    both = (u > infval) & (l < -infval)
    ref[both] = 0.0
    onlyu = (u > infval) & ~(l < -infval)
    ref[onlyu] = np.where(ref[onlyu] < 0.0, ref[onlyu], 0.0)
    onlyl = ~(u > infval) & (l < -infval)
    ref[onlyl] = np.where(ref[onlyl] > 0.0, ref[onlyl], 0.0)
    assert np.array_equal(vy.get(), ref)
    # in_reccone
    w = np.zeros(n)
    vw = Vec(w)
    assert L.OSQPVectorf_in_reccone(vw.h, vl.h, vu.h, infval, 1e-6) == 1
    w[np.where(u < infval)[0][0]] = 1.0
    vw2 = Vec(w)
    assert L.OSQPVectorf_in_reccone(vw2.h, vl.h, vu.h, infval, 1e-6) == 0
    # bounds type: loose (-1) checked before equality (1), else 0
    t = IVec(n=n)
    changed = L.OSQPVectorf_ew_bounds_type(t.h, vl.h, vu.h, 1e-4, infval)
    ref_t = np.where((l < -infval) & (u > infval), -1, np.where(u - l < 1e-4, 1, 0))
    assert changed == 1
    assert np.array_equal(t.get(), ref_t)
    assert L.OSQPVectorf_ew_bounds_type(t.h, vl.h, vu.h, 1e-4, infval) == 0
    vr = Vec(n=n)
    L.OSQPVectorf_set_scalar_conditional(vr.h, t.h, 1e-6, 0.1, 100.0)
    assert np.array_equal(vr.get(), np.where(ref_t == 0, 0.1, np.where(ref_t > 0, 100.0, 1e-6)))
    for v in (vy, vl, vu, vw, vw2, t, vr):
        v.free()


def test_views_and_raw_io(algebra):
    L = algebra.lib()
    a = np.arange(10.0)
    va = Vec(a)
    view = L.OSQPVectorf_view(va.h, 3, 4)
    assert L.OSQPVectorf_length(view) == 4
    out = np.zeros(4)
    L.OSQPVectorf_to_raw(out.ctypes.data, view)
    assert np.array_equal(out, a[3:7])
    L.OSQPVectorf_set_scalar(view, -1.0)
    b = va.get()
    assert np.array_equal(b, np.r_[a[:3], -np.ones(4), a[7:]])
    L.OSQPVectorf_view_free(view)
    c = L.OSQPVectorf_copy_new(va.h)
    out = np.zeros(10)
    L.OSQPVectorf_to_raw(out.ctypes.data, c)
    assert np.array_equal(out, b)
    L.OSQPVectorf_free(c)
    va.free()


def _rand_mats(seed, m, n, d):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=d, format="csc", random_state=rng, data_rvs=rng.standard_normal)
    M = sp.random(n, n, density=d, format="csc", random_state=rng, data_rvs=rng.standard_normal)
    Pf = (M + M.T + sp.eye(n)).tocsc()
    return A, Pf, rng


def _bound(Mabs, xabs, rowlen):
    return 4.0 * EPS * np.maximum(rowlen, 1) * (Mabs @ xabs) + 1e-300


# the last case has 1.2e6 nonzeros: A x and A' y go through the matrix's FP64
# column panels (products of matrices with >= 2^20 nonzeros)
@pytest.mark.parametrize("m,n,d", [(50, 30, 0.2), (3000, 800, 0.01), (20000, 2000, 0.005), (200000, 20000, 3e-4)])
def test_spmv(algebra, m, n, d):
    A, Pf, rng = _rand_mats(m + n, m, n, d)
    L = algebra.lib()
    MA, MP = Mat(A), Mat(Pf, is_triu=True)
    x, y0 = rng.standard_normal(n), rng.standard_normal(m)
    vx, vy = Vec(x), Vec(y0)
    L.OSQPMatrix_Axpy(MA.h, vx.h, vy.h, 2.0, -0.5)
    ref = 2.0 * (A @ x) - 0.5 * y0
    rl = np.diff(A.tocsr().indptr)
    bnd = _bound(abs(A), np.abs(x), rl) * 2 + 4 * EPS * np.abs(0.5 * y0)
    assert np.all(np.abs(vy.get() - ref) <= bnd)
    yy = rng.standard_normal(m)
    vyy, vo = Vec(yy), Vec(n=n)
    L.OSQPMatrix_Atxpy(MA.h, vyy.h, vo.h, 1.0, 0.0)
    assert np.all(np.abs(vo.get() - A.T @ yy) <= _bound(abs(A.T), np.abs(yy), np.diff(A.indptr)))
    vp = Vec(n=n)
    L.OSQPMatrix_Axpy(MP.h, vx.h, vp.h, 1.0, 0.0)
    assert np.all(np.abs(vp.get() - Pf @ x) <= _bound(abs(Pf), np.abs(x), np.diff(Pf.tocsr().indptr)))
    # beta = 0 must overwrite garbage (NaN) in y
    vn = Vec(np.full(m, np.nan))
    L.OSQPMatrix_Axpy(MA.h, vx.h, vn.h, 1.0, 0.0)
    assert np.all(np.isfinite(vn.get()))
    for v in (vx, vy, vyy, vo, vp, vn, MA, MP):
        v.free()


def test_panel_spmv_follows_value_updates(algebra):
    """A large operator (panel path) after update_values and diagonal scaling:
    the panel copies are refreshed before the next product."""
    m, n = 200000, 20000
    A, _, rng = _rand_mats(7, m, n, 3e-4)
    L = algebra.lib()
    MA = Mat(A)
    x, yv = rng.standard_normal(n), rng.standard_normal(m)
    vx, vy, vyv, vo = Vec(x), Vec(n=m), Vec(yv), Vec(n=n)
    L.OSQPMatrix_Axpy(MA.h, vx.h, vy.h, 1.0, 0.0)
    np.testing.assert_allclose(vy.get(), A @ x, rtol=1e-12, atol=1e-12)
    idx = np.array([0, 11, A.nnz - 1], dtype=np.int32)
    newv = np.array([7.0, -3.0, 0.25])
    L.OSQPMatrix_update_values(MA.h, newv.ctypes.data, idx.ctypes.data, 3)
    A2 = A.copy()
    A2.data[idx] = newv
    L.OSQPMatrix_Axpy(MA.h, vx.h, vy.h, 1.0, 0.0)
    np.testing.assert_allclose(vy.get(), A2 @ x, rtol=1e-12, atol=1e-12)
    dl, dr = rng.uniform(0.5, 2.0, m), rng.uniform(0.5, 2.0, n)
    vdl, vdr = Vec(dl), Vec(dr)
    L.OSQPMatrix_lmult_diag(MA.h, vdl.h)
    L.OSQPMatrix_rmult_diag(MA.h, vdr.h)
    A3 = sp.diags(dl) @ A2 @ sp.diags(dr)
    L.OSQPMatrix_Atxpy(MA.h, vyv.h, vo.h, 1.0, 0.0)
    np.testing.assert_allclose(vo.get(), A3.T @ yv, rtol=1e-11, atol=1e-11)
    L.OSQPMatrix_Axpy(MA.h, vx.h, vy.h, 1.0, 0.0)
    np.testing.assert_allclose(vy.get(), A3 @ x, rtol=1e-11, atol=1e-11)
    for v in (MA, vx, vy, vyv, vo, vdl, vdr):
        v.free()


def test_empty_matrices(algebra):
    """tests/lin_alg/test_mat_vec.cpp:71-149: zero-nnz and zero-row matrices."""
    L = algebra.lib()
    A0 = sp.csc_matrix((5, 4))
    MA = Mat(A0)
    vx, vy = Vec(np.ones(4)), Vec(np.full(5, 3.0))
    L.OSQPMatrix_Axpy(MA.h, vx.h, vy.h, 1.0, 2.0)
    assert np.array_equal(vy.get(), np.full(5, 6.0))
    L.OSQPMatrix_Axpy(MA.h, vx.h, vy.h, 1.0, 0.0)
    assert np.array_equal(vy.get(), np.zeros(5))
    Am = sp.csc_matrix((0, 4))
    Mm = Mat(Am)
    vz = Vec(np.full(4, 1.5))
    L.OSQPMatrix_Atxpy(Mm.h, Vec(n=0).h, vz.h, 1.0, 0.0)
    assert np.array_equal(vz.get(), np.zeros(4))
    for v in (MA, vx, vy, Mm, vz):
        v.free()


def test_norms_scaling_and_diag_bitexact(algebra):
    """Ruiz building blocks: col/row inf-norms, l/r diagonal scaling, P triu semantics."""
    O = load_oracle()
    A, Pf, rng = _rand_mats(11, 400, 150, 0.05)
    P = sp.triu(Pf, format="csc")
    P.sort_indices()
    A.sort_indices()
    L = algebra.lib()
    MA, MP = Mat(A), Mat(P, is_triu=True)
    n, m = 150, 400
    E = Vec(n=n)
    L.OSQPMatrix_col_norm_inf(MA.h, E.h)
    assert np.array_equal(E.get(), np.abs(A).max(axis=0).toarray().ravel())
    Er = Vec(n=m)
    L.OSQPMatrix_row_norm_inf(MA.h, Er.h)
    assert np.array_equal(Er.get(), np.abs(A).max(axis=1).toarray().ravel())
    # col_norm_inf of P keeps the builtin meaning: stored triu column
    L.OSQPMatrix_col_norm_inf(MP.h, E.h)
    assert np.array_equal(E.get(), np.abs(P).max(axis=0).toarray().ravel())
    # scaling: l/r multiply, then compare the values through A x with unit vectors
    dl, dr = rng.random(m) + 0.5, rng.random(n) + 0.5
    vdl, vdr = Vec(dl), Vec(dr)
    L.OSQPMatrix_lmult_diag(MA.h, vdl.h)
    L.OSQPMatrix_rmult_diag(MA.h, vdr.h)
    As = A.copy()
    # reference order: x *= d[row] then x *= d[col] (csc_math.c:58-88)
    As.data = (As.data * dl[As.indices]) * np.repeat(dr, np.diff(As.indptr))
    L.OSQPMatrix_row_norm_inf(MA.h, Er.h)
    assert np.array_equal(Er.get(), np.abs(As).max(axis=1).toarray().ravel())
    L.OSQPMatrix_col_norm_inf(MA.h, E.h)
    assert np.array_equal(E.get(), np.abs(As).max(axis=0).toarray().ravel())
    dp = rng.random(n) + 0.5
    vdp = Vec(dp)
    L.OSQPMatrix_lmult_diag(MP.h, vdp.h)
    L.OSQPMatrix_rmult_diag(MP.h, vdp.h)
    L.OSQPMatrix_mult_scalar(MP.h, 0.37)
    Ps = P.copy()
    Ps.data = ((Ps.data * dp[Ps.indices]) * np.repeat(dp, np.diff(Ps.indptr))) * 0.37
    L.OSQPMatrix_col_norm_inf(MP.h, E.h)
    assert np.array_equal(E.get(), np.abs(Ps).max(axis=0).toarray().ravel())
    # the full copy stays exactly symmetric: row norms equal the symmetric ones
    Eo = np.zeros(n)
    # keep the int32 copies alive across the call (ctypes takes raw addresses)
    Pp32, Pi32, Px64 = Ps.indptr.astype(np.int32), Ps.indices.astype(np.int32), np.ascontiguousarray(Ps.data)
    O.oracle_csc_row_norm_inf_sym_triu(n, Pp32.ctypes.data, Pi32.ctypes.data, Px64.ctypes.data, Eo.ctypes.data)
    L.OSQPMatrix_row_norm_inf(MP.h, E.h)
    assert np.array_equal(E.get(), Eo)
    dg = Vec(n=n)
    L.OSQPMatrix_extract_diag(MP.h, dg.h)
    assert np.array_equal(dg.get(), Ps.diagonal())
    Dw = rng.random(m)
    vD, out = Vec(Dw), Vec(n=n)
    L.OSQPMatrix_AtDA_extract_diag(MA.h, vD.h, out.h)
    ref = (As.T @ sp.diags(Dw) @ As).diagonal()
    np.testing.assert_allclose(out.get(), ref, rtol=64 * EPS * 10, atol=1e-300)
    for v in (MA, MP, E, Er, vdl, vdr, vdp, dg, vD, out):
        v.free()


def test_update_values_and_submatrix(algebra):
    A, Pf, rng = _rand_mats(21, 60, 25, 0.2)
    P = sp.triu(Pf, format="csc")
    L = algebra.lib()
    MA, MP = Mat(A), Mat(P, is_triu=True)
    assert L.OSQPMatrix_get_nz(MA.h) == A.nnz and L.OSQPMatrix_get_nz(MP.h) == P.nnz
    assert L.OSQPMatrix_get_m(MA.h) == 60 and L.OSQPMatrix_get_n(MA.h) == 25
    idx = np.array([0, 3, 5], dtype=np.int32)
    newv = np.array([9.0, -4.0, 2.5])
    L.OSQPMatrix_update_values(MA.h, newv.ctypes.data, idx.ctypes.data, 3)
    A2 = A.copy()
    A2.data[idx] = newv
    x = rng.standard_normal(25)
    vx, vy = Vec(x), Vec(n=60)
    L.OSQPMatrix_Axpy(MA.h, vx.h, vy.h, 1.0, 0.0)
    np.testing.assert_allclose(vy.get(), A2 @ x, rtol=1e-12, atol=1e-12)
    yv = rng.standard_normal(60)
    vyv, vo = Vec(yv), Vec(n=25)
    L.OSQPMatrix_Atxpy(MA.h, vyv.h, vo.h, 1.0, 0.0)
    np.testing.assert_allclose(vo.get(), A2.T @ yv, rtol=1e-12, atol=1e-12)
    pv = rng.standard_normal(P.nnz)
    L.OSQPMatrix_update_values(MP.h, pv.ctypes.data, None, P.nnz)
    P2 = P.copy()
    P2.data = pv
    Pf2 = P2 + sp.triu(P2, 1).T
    vp = Vec(n=25)
    L.OSQPMatrix_Axpy(MP.h, vx.h, vp.h, 1.0, 0.0)
    np.testing.assert_allclose(vp.get(), Pf2 @ x, rtol=1e-12, atol=1e-12)
    rows = IVec((rng.random(60) < 0.4).astype(np.int32))
    sel = rows.get().astype(bool)
    sub = L.OSQPMatrix_submatrix_byrows(MA.h, rows.h)
    assert L.OSQPMatrix_get_m(sub) == sel.sum()
    vs = Vec(n=int(sel.sum()))
    L.OSQPMatrix_Axpy(sub, vx.h, vs.h, 1.0, 0.0)
    np.testing.assert_allclose(vs.get(), (A2 @ x)[sel], rtol=1e-12, atol=1e-12)
    L.OSQPMatrix_free(sub)
    for v in (MA, MP, vx, vy, vyv, vo, vp, rows, vs):
        v.free()


@pytest.mark.parametrize("seed", [0, 1])
def test_spmv_skewed_rows(algebra, seed):
    """Row-block partition: rows far longer than one block's chunk (2048
    nonzeros), empty rows and many tiny rows in one matrix."""
    rng = np.random.default_rng(seed)
    m, n = 3000, 20000
    lens = rng.integers(0, 6, m)
    lens[[5, 1700]] = [9000, 20000]      # long rows (one is fully dense)
    lens[100:140] = 0                    # empty rows
    rows = np.repeat(np.arange(m), lens)
    cols = np.concatenate([rng.choice(n, k, replace=False) for k in lens])
    A = sp.csc_matrix((rng.standard_normal(len(rows)), (rows, cols)), shape=(m, n))
    L = algebra.lib()
    MA = Mat(A)
    x = rng.standard_normal(n)
    vx, vy = Vec(x), Vec(n=m)
    L.OSQPMatrix_Axpy(MA.h, vx.h, vy.h, 1.0, 0.0)
    bnd = _bound(abs(A), np.abs(x), np.diff(A.tocsr().indptr))
    assert np.all(np.abs(vy.get() - A @ x) <= bnd)
    yv = rng.standard_normal(m)
    vyv, vo = Vec(yv), Vec(n=n)
    L.OSQPMatrix_Atxpy(MA.h, vyv.h, vo.h, 1.0, 0.0)
    assert np.all(np.abs(vo.get() - A.T @ yv) <= _bound(abs(A.T), np.abs(yv), np.diff(A.indptr)))
    for v in (MA, vx, vy, vyv, vo):
        v.free()


def test_spmv_row_block_window_edges(algebra):
    """Row-block partition at its limits: rows of 2043..2049 nonzeros (either
    side of the 2045-nonzero block capacity), blocks starting at every
    alignment, and rows of 17..300 nonzeros reduced by lane groups."""
    rng = np.random.default_rng(7)
    n = 4000
    lens = np.concatenate([[2043, 2044, 2045, 2046, 2047, 2048, 2049, 1, 2, 3, 17, 20, 33, 64, 65, 100, 300, 0],
                           rng.integers(0, 120, 700), rng.integers(15, 40, 500), [2045, 1, 2046]])
    m = len(lens)
    rows = np.repeat(np.arange(m), lens)
    cols = np.concatenate([rng.choice(n, k, replace=False) for k in lens])
    A = sp.csc_matrix((rng.standard_normal(len(rows)), (rows, cols)), shape=(m, n))
    L = algebra.lib()
    MA = Mat(A)
    x = rng.standard_normal(n)
    vx, vy = Vec(x), Vec(n=m)
    L.OSQPMatrix_Axpy(MA.h, vx.h, vy.h, 1.0, 0.0)
    assert np.all(np.abs(vy.get() - A @ x) <= _bound(abs(A), np.abs(x), np.diff(A.tocsr().indptr)))
    yv = rng.standard_normal(m)
    vyv, vo = Vec(yv), Vec(n=n)
    L.OSQPMatrix_Atxpy(MA.h, vyv.h, vo.h, 1.0, 0.0)
    assert np.all(np.abs(vo.get() - A.T @ yv) <= _bound(abs(A.T), np.abs(yv), np.diff(A.indptr)))
    for v in (MA, vx, vy, vyv, vo):
        v.free()


def test_panel_spmv_skewed_rows(algebra):
    """The FP64 panel path (over 2^20 nonzeros): rows longer than a task (the
    multi-pass long-row path), a run of empty rows and rows of every length
    from 0 to 120, so task windows start at every alignment; A x and A' y
    against scipy within the summation-order bound."""
    rng = np.random.default_rng(4)
    m, n = 20000, 4000
    lens = rng.integers(0, 121, m)
    lens[[3, 777, 15000]] = [n, 2500, 900]   # longer than a 383-entry task
    lens[5000:5100] = 0                      # empty rows
    rows = np.repeat(np.arange(m), lens)
    cols = np.concatenate([rng.choice(n, k, replace=False) for k in lens])
    A = sp.csc_matrix((rng.standard_normal(len(rows)), (rows, cols)), shape=(m, n))
    assert A.nnz >= 1 << 20
    L = algebra.lib()
    MA = Mat(A)
    for rep in range(2):  # the second product runs on panels built by the first
        x = rng.standard_normal(n)
        vx, vy = Vec(x), Vec(n=m)
        L.OSQPMatrix_Axpy(MA.h, vx.h, vy.h, 1.0, 0.0)
        bnd = _bound(abs(A), np.abs(x), np.diff(A.tocsr().indptr))
        assert np.all(np.abs(vy.get() - A @ x) <= bnd)
        yv = rng.standard_normal(m)
        vyv, vo = Vec(yv), Vec(n=n)
        L.OSQPMatrix_Atxpy(MA.h, vyv.h, vo.h, 1.0, 0.0)
        assert np.all(np.abs(vo.get() - A.T @ yv) <= _bound(abs(A.T), np.abs(yv), np.diff(A.indptr)))
        for v in (vx, vy, vyv, vo):
            v.free()
    MA.free()
