"""The batched serving engine against the CPU oracle, and its result contract.

The engine has no counterpart in the reference, but each problem it solves is
an OSQP problem: its per-problem arithmetic restates the host core's ADMM
step and the PCG linear solver.  So its results are checked here against the
oracle (oracle/osqp_oracle.c, FP64; the three references of
tests/oracle_refs.py), not against the same library's single-problem engine:
  * a seeded MPC stream (nx = 12, nu = 4, N ~ U{10..50}, 256 requests, eps
    1e-4): every status equals the oracle's direct (LDL') path and its PCG
    path; every objective is within 1e-4 relative of the FP64 optimum; x and
    y are within the stated bounds of the oracle's PCG path, the same
    algorithm (FP64 engine: 1e-5; FP32-PCG engine: 1e-3);
  * portfolios of 500, 2000 and 5000 assets with two new mu (osqp_update_data_vec
    semantics, warm-started re-solves), eps 1e-4, against the oracle answers
    of tests/golden/make_serve_fixtures.py: statuses equal both oracle paths';
    each objective is within 1e-4 of the oracle's own accuracy on that
    request (the larger of OSQP's gap-test bound and the oracle PCG path's
    error against the FP64 optimum);
  * a QP with rows of every length class the team kernel's products treat
    apart (short, just over 64 entries, around multiples of 4, 8 and 2048);
  * OSQP's result contract (src/auxil.c:598-674): infeasible problems return
    x = y = OSQP_NAN and the normalised certificate, equal to the
    single-problem engine's; a non-convex problem returns obj = OSQP_NAN;
  * time_limit (src/osqp_api.c:743-759) stops a solve with status 8;
  * osqp_update_data_mat (src/osqp_api.c:1309-1409) on a resident problem
    reaches the reference's held solutions (tests/update_matrices).
"""
import functools
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import OracleSolver
from osqp_amd import OSQP, OSQP_NAN
from osqp_amd import problems as pr
from osqp_amd.serve import BatchEngine

from .conftest import GOLDEN
from .oracle_refs import gap_bound, optimum, rel, snapshot, solve, vec_err

pytestmark = pytest.mark.gpu

ST = dict(eps_abs=1e-4, eps_rel=1e-4, max_iter=4000)
OBJ_TOL = 1e-4
XY_TOL = {64: 1e-5, 32: 1e-3}


def oracle(p, linsys=2, **kw):
    st = dict(ST)
    st.update(kw)
    return solve(p, linsys, **st)


def run_all(eng, probs, resident=False):
    ids = eng.add_many(probs, resident=resident)
    got = {}
    while eng.pending():
        for i in eng.step():
            got[i] = eng.poll(i)
    return ids, [got.get(i) for i in ids]


def mpc_stream(k, seed=7):
    rng = np.random.default_rng(seed)
    return [pr.mpc(int(N), seed=1000 * seed + j) for j, N in enumerate(rng.integers(10, 51, k))]


@functools.lru_cache(maxsize=None)
def mpc_refs():
    """The stream and, per request, the oracle's PCG and direct answers and the optimum."""
    probs = mpc_stream(256)
    return probs, [(snapshot(oracle(p, 2)), snapshot(oracle(p, 1)), optimum(p, **ST)) for p in probs]


@pytest.mark.parametrize("precision", [32, 64])
def test_mpc_stream_vs_oracle(precision):
    probs, refs = mpc_refs()
    eng = BatchEngine(precision=precision, **ST)
    _, res = run_all(eng, probs)
    eng.close()
    worst = dict(obj_opt=0.0, x_pcg=0.0, y_pcg=0.0, obj_pcg=0.0, direct_obj_opt=0.0)
    for k, (r, (pc, di, op)) in enumerate(zip(res, refs)):
        assert r.info.status_val == pc["status"] == di["status"], (k, r.info.status, pc["status"], di["status"])
        if di["status"] != 1:
            continue
        worst["obj_opt"] = max(worst["obj_opt"], rel(r.info.obj_val, op["obj"]))
        worst["direct_obj_opt"] = max(worst["direct_obj_opt"], rel(di["obj"], op["obj"]))
        worst["obj_pcg"] = max(worst["obj_pcg"], rel(r.info.obj_val, pc["obj"]))
        worst["x_pcg"] = max(worst["x_pcg"], vec_err(r.x, pc["x"]))
        worst["y_pcg"] = max(worst["y_pcg"], vec_err(r.y, pc["y"]))
    print(f"batch{precision} MPC vs oracle: {worst}")
    assert worst["obj_opt"] <= OBJ_TOL
    assert worst["x_pcg"] <= XY_TOL[precision] and worst["y_pcg"] <= XY_TOL[precision]


def portfolio_refs(n):
    """The stream of tests/golden/make_serve_fixtures.py and its oracle answers."""
    p = pr.portfolio(n, seed=40 + n)
    rng = np.random.default_rng(n)
    qs = [p["q"]] + [pr.portfolio_q(rng.standard_normal(n), p["k"]) for _ in range(2)]
    z = np.load(os.path.join(GOLDEN, "portfolio_oracle.npz"))
    refs = []
    for j, q in enumerate(qs):
        d_st, d_obj, c_st, c_obj, o_obj = z[f"n{n}_s{j}_info"]
        x, y = z[f"n{n}_s{j}_x"], z[f"n{n}_s{j}_y"]
        refs.append(dict(direct=(int(d_st), d_obj), pcg=(int(c_st), c_obj), opt=(o_obj, x, y),
                         bound=gap_bound(p, x, y, 1e-4, 1e-4, q)))
    return p, qs, refs


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("n", [500, 2000, 5000])
def test_portfolio_updates_vs_oracle(n, precision):
    """A resident portfolio re-solved after two new mu (osqp_update_data_vec
    + warm-started osqp_solve), max_iter 10000 so no solve ends on the
    iteration limit.  Statuses equal the oracle's direct path.  The objective
    is held to the oracle's own accuracy on the same request: within 1e-4
    (relative) of the larger of OSQP's gap-test bound and the error of the
    oracle's PCG path -- the algorithm the engine runs -- against the
    optimum.  (On these problems both oracle paths land 1e-4 to 5e-4 from the
    optimum at eps 1e-4; a warm re-solve of the PCG path can stop early, as
    the engine's does, where the termination test is met by a point 1e-2 away
    in x.)"""
    p, qs, refs = portfolio_refs(n)
    eng = BatchEngine(precision=precision, **dict(ST, max_iter=10000))
    pid = eng.add(p["P"], p["q"], p["A"], p["l"], p["u"], resident=True)
    got = []
    for j in range(len(qs)):
        if j:
            eng.resolve(pid, qs[j])
        while eng.pending():
            eng.step()
        got.append(eng.poll(pid))
    eng.close()
    for j, (r, ref) in enumerate(zip(got, refs)):
        f_opt, x_opt, y_opt = ref["opt"]
        e_gpu, e_dir, e_pcg = rel(r.info.obj_val, f_opt), rel(ref["direct"][1], f_opt), rel(ref["pcg"][1], f_opt)
        print(f"portfolio{n} batch{precision} solve {j}: {r.info.status} it={r.info.iter} obj vs optimum "
              f"{e_gpu:.2e} (oracle direct {e_dir:.2e}, oracle PCG {e_pcg:.2e}, gap bound {ref['bound']:.2e}) "
              f"x {vec_err(r.x, x_opt):.2e} y {vec_err(r.y, y_opt):.2e}")
        assert r.info.status_val == ref["direct"][0] == ref["pcg"][0] == 1
        assert e_gpu <= max(ref["bound"], e_pcg) + OBJ_TOL


def long_row_qp(seed=11):
    """A QP whose A has rows of every length class the team kernel treats
    apart: short rows (row groups) and rows longer than 64 entries, just
    above it and around multiples of 4, 8, 64 and 2048 (the FP32 engine's
    compact long rows: 4-entry chunks per lane, rows padded to 8, rounds of
    2048 entries per wavefront), and one of 1000."""
    rng = np.random.default_rng(seed)
    n = 3000
    lens = [65, 66, 67, 69, 71, 72, 73, 127, 128, 129, 255, 256, 257, 1000, 2047, 2048, 2049]
    lens += list(rng.integers(1, 21, 500))
    rows, cols, vals = [], [], []
    for i, L in enumerate(lens):
        c = rng.choice(n, int(L), replace=False)
        rows += [i] * int(L)
        cols += list(c)
        vals += list(rng.standard_normal(int(L)) / np.sqrt(L))
    A = sp.csc_matrix((vals, (rows, cols)), shape=(len(lens), n))
    P = sp.triu(sp.diags(rng.random(n) + 0.1), format="csc")
    m = A.shape[0]
    return dict(P=P, q=rng.standard_normal(n), A=A, l=-0.5 - rng.random(m), u=0.5 + rng.random(m))


@pytest.mark.parametrize("precision", [32, 64])
def test_long_row_lengths_vs_oracle(precision):
    """Every row-length class of the team's sparse products against the
    oracle: statuses equal both oracle paths', the objective within 1e-4 of
    the FP64 optimum's accuracy bound, x and y within 1e-3 of the oracle's
    PCG path for both engines.  (Here the FP64 engine and the oracle's PCG
    path take the same 700 ADMM iterations but part by about 1e-4 in y: each
    CG solve stops once its residual is under the tolerance, and the GPU's
    blocked dot products move that stop within it; over 700 inexact solves
    the points differ at the tolerance's level, not at rounding's.)"""
    p = long_row_qp()
    pc, di, op = oracle(p, 2), oracle(p, 1), optimum(p, **ST)
    eng = BatchEngine(precision=precision, **ST)
    _, (r,) = run_all(eng, [p])
    eng.close()
    e_obj = rel(r.info.obj_val, op["obj"])
    gb = gap_bound(p, op["x"], op["y"], 1e-4, 1e-4)
    e_x, e_y = vec_err(r.x, pc.x), vec_err(r.y, pc.y)
    print(f"long rows batch{precision}: {r.info.status} it {r.info.iter} (oracle PCG {pc.info.status} "
          f"it {pc.info.iter}) obj vs optimum {e_obj:.1e} (gap bound {gb:.1e}) x {e_x:.1e} y {e_y:.1e}")
    assert r.info.status_val == pc.info.status_val == di.info.status_val == 1
    assert e_obj <= max(gb, rel(pc.info.obj_val, op["obj"])) + OBJ_TOL
    assert e_x <= 1e-3 and e_y <= 1e-3


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("case", [1, 2, 3, 4])
def test_infeasibility_contract(case, precision):
    """Statuses, OSQP_NAN fills and certificates equal the single-problem
    engine's (src/auxil.c:598-674)."""
    p = pr.primal_dual_infeasibility(case)
    eng = BatchEngine(precision=precision, **ST)
    _, (rb,) = run_all(eng, [p])
    eng.close()
    s = OSQP(precision=precision)
    assert s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **ST) == 0
    rs = s.solve()
    s.close()
    assert rb.info.status == p["status"] == rs.info.status
    if rs.info.status_val in (1, 2):
        assert np.all(rb.prim_inf_cert == OSQP_NAN) and np.all(rb.dual_inf_cert == OSQP_NAN)
        return
    assert np.all(rb.x == OSQP_NAN) and np.all(rb.y == OSQP_NAN)
    if "primal" in p["status"]:
        cert, other = rb.prim_inf_cert, rb.dual_inf_cert
        ref = rs.prim_inf_cert
    else:
        cert, other = rb.dual_inf_cert, rb.prim_inf_cert
        ref = rs.dual_inf_cert
    assert abs(np.max(np.abs(cert)) - 1.0) < 1e-12
    assert np.all(other == OSQP_NAN)
    np.testing.assert_allclose(cert, ref, rtol=0, atol=1e-6)


@pytest.mark.parametrize("precision", [64, 32])
def test_non_convex_obj_is_osqp_nan(precision):
    """tests/non_cvx/test_non_cvx.cpp:47-71 (the reference test fixture's
    settings, tests/osqp_tester.h:71-81, with sigma = 5 and adaptive rho off):
    diverging residuals end with status non-convex (src/auxil.c:831-840),
    obj_val = OSQP_NAN (2143289344.0) and x = y = OSQP_NAN, as in the
    single-problem engine."""
    p = pr.non_cvx()
    st = dict(rho=0.1, alpha=1.6, max_iter=2000, scaling=1, eps_abs=1e-5, eps_rel=1e-5, adaptive_rho=0, sigma=5.0)
    eng = BatchEngine(precision=precision, **st)
    _, (r,) = run_all(eng, [p])
    eng.close()
    s = OSQP(precision=precision)
    assert s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **st) == 0
    rs = s.solve()
    s.close()
    assert r.info.status_val == rs.info.status_val == 9
    assert r.info.obj_val == OSQP_NAN
    assert np.all(r.x == OSQP_NAN) and np.all(r.y == OSQP_NAN)


def test_time_limit_in_batch():
    p = pr.portfolio(2000, seed=3)
    eng = BatchEngine(precision=64, eps_abs=1e-12, eps_rel=1e-12, max_iter=10**8, check_termination=0,
                      time_limit=0.05)
    _, (r,) = run_all(eng, [p])
    eng.close()
    assert r.info.status_val == 8, r.info.status
    assert 0 < r.info.iter < 10**8
    # a problem that ends before the limit is unaffected
    q = pr.mpc(10, seed=1)
    eng = BatchEngine(precision=64, time_limit=30.0, **ST)
    _, (r1,) = run_all(eng, [q])
    eng.close()
    eng = BatchEngine(precision=64, **ST)
    _, (r2,) = run_all(eng, [q])
    eng.close()
    assert r1.info.status_val == r2.info.status_val == 1
    assert r1.info.iter == r2.info.iter and r1.info.obj_val == r2.info.obj_val


@pytest.mark.parametrize("mode", [2, 3])
def test_adaptive_rho_time_and_kkt_modes(mode):
    """The time-based and KKT-error-based rho adaptation (src/osqp_api.c:
    776-877) run in the batch and reach the oracle's status and objective
    (the oracle runs the same mode; the time mode's interval depends on
    timing, so only the answer is compared)."""
    probs = mpc_stream(8, seed=3) + [pr.portfolio(500, seed=9)]
    eng = BatchEngine(precision=64, adaptive_rho=mode, **ST)
    _, res = run_all(eng, probs)
    eng.close()
    for p, r in zip(probs, res):
        o = oracle(p, adaptive_rho=mode)
        assert r.info.status_val == o.info.status_val == 1
        assert rel(r.info.obj_val, o.info.obj_val) <= OBJ_TOL


SECTIONS = [("P", "idx"), ("P", "all"), ("A", "idx"), ("A", "all"), ("PA", "idx"), ("PA", "all")]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("which,mode", SECTIONS)
def test_update_data_mat_resident(which, mode, precision):
    """tests/update_matrices/test_update_matrices.cpp:138-384 on a resident
    problem: new P and / or A values, then a re-solve, reach the held
    solutions (TESTS_TOL 1e-4), and equal the oracle after the same update."""
    p = pr.update_matrices_problem()
    st = dict(rho=0.1, alpha=1.6, max_iter=1000, scaling=1, eps_abs=1e-5, eps_rel=1e-5)
    eng = BatchEngine(precision=precision, **st)
    pid = eng.add(p["P"], p["q"], p["A"], p["l"], p["u"], resident=True)
    while eng.pending():
        eng.step()
    r0 = eng.poll(pid)
    assert r0.info.status_val == 1
    assert np.max(np.abs(r0.x - p["x"])) < 1e-4
    kw = {}
    if "P" in which:
        kw["Px"] = p["Pu_new"].data
        if mode == "idx":
            kw["Px_idx"] = np.arange(p["Pu_new"].nnz)
    if "A" in which:
        kw["Ax"] = p["A_new"].data
        if mode == "idx":
            kw["Ax_idx"] = np.arange(p["A_new"].nnz)
    eng.update_mat(pid, **kw)
    eng.resolve(pid)
    while eng.pending():
        eng.step()
    r = eng.poll(pid)
    eng.close()
    key = {"P": "P_new", "A": "A_new", "PA": "PA_new"}[which]
    assert r.info.status_val == 1
    assert np.max(np.abs(r.x - p["x_" + key])) < 1e-4
    assert np.max(np.abs(r.y - p["y"])) < 1e-4
    assert abs(r.info.obj_val - p["obj_" + key]) < 1e-4
    o = OracleSolver()
    assert o.setup(p["P"], p["q"], p["A"], p["l"], p["u"], linsys_solver=2, **st) == 0
    o.solve()
    assert o.update_data_mat(**kw) == 0
    o.solve()
    assert o.info.status_val == 1
    assert abs(r.info.obj_val - o.info.obj_val) <= 1e-4 * max(1.0, abs(o.info.obj_val))


def test_update_data_mat_partial_and_rejected():
    """A partial update (a few indexed values) equals a fresh problem with
    those values; bad counts and indices are rejected (1: P, 2: A) and leave
    the problem as it was."""
    p = pr.portfolio(300, seed=12)
    A2 = p["A"].copy()
    idx = np.arange(0, A2.nnz, 7)
    A2.data[idx] = A2.data[idx] * 1.5
    eng = BatchEngine(precision=64, **ST)
    pid = eng.add(p["P"], p["q"], p["A"], p["l"], p["u"], resident=True)
    while eng.pending():
        eng.step()
    eng.poll(pid)
    with pytest.raises(ValueError):
        eng.update_mat(pid, Ax=np.ones(3))  # partial without indices
    with pytest.raises(ValueError):
        eng.update_mat(pid, Ax=np.ones(2), Ax_idx=np.array([0, A2.nnz]))  # index out of range
    with pytest.raises(ValueError):
        eng.update_mat(pid, Px=np.ones(p["P"].nnz + 1))
    eng.update_mat(pid, Ax=A2.data[idx], Ax_idx=idx)
    eng.resolve(pid)
    while eng.pending():
        eng.step()
    r = eng.poll(pid)
    eng.close()
    o = oracle(dict(p, A=A2))
    assert r.info.status_val == o.info.status_val == 1
    assert rel(r.info.obj_val, o.info.obj_val) <= OBJ_TOL


class DeviceArrays:
    """Device copies of host arrays, made with the HIP runtime's own hipMalloc
    (no framework in between), freed with hipFree."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def put(self, a):
        C = self.C
        a = np.ascontiguousarray(a)
        ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(ptr), C.c_size_t(max(a.nbytes, 8))) == 0
        assert self.hip.hipMemcpy(ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # H2D
        self.ptrs.append(ptr)
        return ptr.value

    def free(self):
        for ptr in self.ptrs:
            self.hip.hipFree(ptr)
        self.ptrs = []


def test_device_pointer_ingestion():
    """Problem arrays in device memory are taken as osqp_batch_add and
    osqp_setup take host arrays: same results, bit for bit."""
    import ctypes as C

    from osqp_amd import _types as T
    from osqp_amd.serve import _bind
    from osqp_amd.solver import load_library
    p = pr.mpc(15, seed=8)
    Ph, Ah = T.CscHolder(p["P"]), T.CscHolder(p["A"])
    lib = _bind(load_library())
    eng = BatchEngine(precision=64, **ST)  # initialises the backend before any hipMalloc here
    D = DeviceArrays()
    arrays = dict(Pp=D.put(Ph.p), Pi=D.put(Ph.i), Px=D.put(Ph.x), Ap=D.put(Ah.p), Ai=D.put(Ah.i), Ax=D.put(Ah.x),
                  q=D.put(p["q"]), l=D.put(np.clip(p["l"], -1e30, 1e30)), u=D.put(np.clip(p["u"], -1e30, 1e30)))
    Pd = T.OSQPCscMatrix(Ph.m, Ph.n, arrays["Pp"], arrays["Pi"], arrays["Px"], int(p["P"].nnz), -1, 0)
    Ad = T.OSQPCscMatrix(Ah.m, Ah.n, arrays["Ap"], arrays["Ai"], arrays["Ax"], int(p["A"].nnz), -1, 0)
    pid = C.c_int()
    assert lib.osqp_batch_add(eng._b, C.addressof(Pd), arrays["q"], C.addressof(Ad),
                              arrays["l"], arrays["u"], None, None, 0, C.byref(pid)) == 0
    eng._dims[pid.value] = (Ph.n, Ah.m)
    got = {}
    while eng.pending():
        for i in eng.step():
            got[i] = eng.poll(i)
    rd = got[pid.value]
    _, (rh,) = run_all(eng, [p])
    eng.close()
    assert rd.info.status_val == rh.info.status_val == 1
    assert rd.info.obj_val == rh.info.obj_val
    np.testing.assert_array_equal(rd.x, rh.x)
    # osqp_setup with the same device arrays
    s = OSQP(precision=64)
    s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **ST)
    rs_host = s.solve()
    settings = T.default_settings()
    T.apply_settings(settings, **ST)
    solver = C.POINTER(T.OSQPSolver)()
    assert lib.osqp_setup(C.byref(solver), C.addressof(Pd), arrays["q"], C.addressof(Ad),
                          arrays["l"], arrays["u"], Ah.m, Ph.n, C.addressof(settings)) == 0
    lib.osqp_solve(solver)
    info = solver.contents.info.contents
    x = np.ctypeslib.as_array(solver.contents.solution.contents.x, shape=(Ph.n,)).copy()
    assert info.status_val == 1 and info.obj_val == rs_host.info.obj_val
    np.testing.assert_array_equal(x, rs_host.x)
    lib.osqp_cleanup(solver)
    s.close()
    D.free()
