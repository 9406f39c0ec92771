"""Solver-level parity of the HIP engine (osqp_setup/osqp_solve through the C ABI).

The reference's solver tests (tests/basic_qp*, basic_lp, unconstrained,
no_active_set, primal_dual_infeasibility, primal_infeasibility, non_cvx,
large_qp, update_matrices) are restated here against the closed-form answers
and statuses those tests hold, at the reference test fixture's settings
(tests/osqp_tester.h:71-81) and its TESTS_TOL = 1e-4 for x, y and the
objective -- for both arithmetics: precision=64 (plain FP64 backend) and
precision=32 (FP32-PCG with FP64 defect correction, residuals and checks).

Problems without a closed form are compared with the oracle
(oracle/osqp_oracle.c, FP64; tests/oracle_refs.py):
  * the status equals the oracle's direct (LDL') path, the reference's default;
  * the objective is held to the reference's own accuracy plus 1e-4: with
    f* the FP64 optimum, |f - f*| <= (max(B, |f_pcg - f*|) + 1e-4) max(1, |f*|),
    B being OSQP's duality-gap stopping bound (tests/oracle_refs.py gap_bound)
    and f_pcg the oracle's PCG path -- the algorithm the engines run -- on the
    same problem.  (At eps 1e-4 both oracle paths can land 1e-4 to 5e-4 from
    the optimum; where they are within 1e-4, the engines are too.)
  * the point meets the eps test on its own (FP64 KKT certificate with scipy);
  * the FP64 engine follows the oracle's PCG trajectory: the same ADMM
    iteration count and rho updates and the same point to rounding
    (test_fp64_trajectory_matches_oracle_pcg).
"""
import os

import numpy as np
import pytest

from oracle import OracleSolver
from osqp_amd import OSQP, OSQP_NAN
from osqp_amd import problems as pr

from .conftest import GOLDEN
from .oracle_refs import gap_bound, optimum, rel, snapshot, solve, vec_err

pytestmark = pytest.mark.gpu

FIXTURE = dict(rho=0.1, alpha=1.6, max_iter=2000, scaling=1, eps_abs=1e-5, eps_rel=1e-5)
PRECISIONS = [64, 32]


def tol_for(precision):
    """TESTS_TOL (tests/osqp_tester.h:12-16): 1e-4 for both engines, whose
    termination tests and iterates outside the CG are FP64."""
    return 1e-4


def gpu_solve(prob, precision=64, **kw):
    s = OSQP(precision=precision)
    st = dict(FIXTURE)
    st.update(kw)
    rc = s.setup(prob["P"], prob["q"], prob["A"], prob["l"], prob["u"], **st)
    assert rc == 0, f"setup failed rc={rc}"
    res = s.solve()
    return s, res


OBJ_TOL = 1e-4


def check_against_oracle(p, r, st, precision, name=""):
    """Status vs the oracle's direct path; objective vs the FP64 optimum within
    1e-4 beyond the oracle PCG path's own miss; prints the errors."""
    di = snapshot(solve(p, 1, **st))
    pc = snapshot(solve(p, 2, **st))
    assert r.info.status_val == di["status"], (name, r.info.status, di["status"])
    op = optimum(p, **st)
    if di["status"] != 1 or op is None:
        print(f"{name} p{precision}: status {r.info.status_val} (oracle {di['status']}), no optimum to compare")
        return di, pc, op
    e_gpu, e_dir, e_pcg = rel(r.info.obj_val, op["obj"]), rel(di["obj"], op["obj"]), rel(pc["obj"], op["obj"])
    gb = gap_bound(p, op["x"], op["y"], st.get("eps_abs", 1e-3), st.get("eps_rel", 1e-3))
    print(f"{name} p{precision}: it {r.info.iter} (pcg {pc['iter']}, direct {di['iter']}) obj vs optimum {e_gpu:.1e} "
          f"(oracle direct {e_dir:.1e}, PCG {e_pcg:.1e}, gap bound {gb:.1e}) x {vec_err(r.x, op['x']):.1e} "
          f"y {vec_err(r.y, op['y']):.1e}")
    assert e_gpu <= max(gb, e_pcg) + OBJ_TOL, (name, e_gpu, e_pcg, gb)
    return di, pc, op


def oracle_solve(prob, **kw):
    o = OracleSolver()
    st = dict(FIXTURE)
    st.update(kw)
    assert o.setup(prob["P"], prob["q"], prob["A"], prob["l"], prob["u"], linsys_solver=1, **st) == 0
    o.solve()
    return o


@pytest.mark.parametrize("precision", PRECISIONS)
def test_basic_qp(precision):
    p = pr.basic_qp()
    s, r = gpu_solve(p, precision)
    T = tol_for(precision)
    assert r.info.status_val == 1, r.info.status
    assert np.max(np.abs(r.x - p["x"])) < T
    assert np.max(np.abs(r.y - p["y"])) < T
    assert abs(r.info.obj_val - p["obj"]) < T
    # prim/dual infeasibility certificates are OSQP_NAN for a solved problem
    assert np.all(r.prim_inf_cert == OSQP_NAN) and np.all(r.dual_inf_cert == OSQP_NAN)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_basic_qp_update_vectors(precision):
    """test_basic_qp.cpp:535-716: update q, then l/u, re-solve."""
    p = pr.basic_qp()
    s, r = gpu_solve(p, precision)
    o = OracleSolver()
    o.setup(p["P"], p["q"], p["A"], p["l"], p["u"], linsys_solver=1, **FIXTURE)
    o.solve()
    for upd in (dict(q=p["q_new"]), dict(l=p["l_new"]), dict(u=p["u_new"]), dict(l=p["l_new"], u=p["u_new"])):
        assert s.update(**upd) == 0
        r = s.solve()
        o.update_data_vec(**upd)
        o.solve()
        assert r.info.status_val == o.info.status_val
        if o.info.status_val == 1:
            assert abs(r.info.obj_val - o.info.obj_val) <= tol_for(precision) * max(1.0, abs(o.info.obj_val))
    # l > u is rejected and leaves the data unchanged
    assert s.update(l=np.array([2.0, 0, 0, 0]), u=np.array([1.0, 1, 1, 1])) == 1


@pytest.mark.parametrize("precision", PRECISIONS)
def test_basic_qp2_update(precision):
    p = pr.basic_qp2()
    s, r = gpu_solve(p, precision, eps_abs=1e-6, eps_rel=1e-6)
    T = tol_for(precision)
    assert r.info.status_val == 1
    assert np.max(np.abs(r.x - p["x"])) / np.max(np.abs(p["x"])) < T
    assert abs(r.info.obj_val - p["obj"]) / abs(p["obj"]) < T
    assert s.update(q=p["q_new"], u=p["u_new"]) == 0
    r = s.solve()
    assert r.info.status_val == 1
    assert np.max(np.abs(r.x - p["x_new"])) / np.max(np.abs(p["x_new"])) < T
    assert abs(r.info.obj_val - p["obj_new"]) / abs(p["obj_new"]) < T


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("polishing,expected", [(0, 0), (1, 1)])
def test_basic_qp2_polish(polishing, expected, precision):
    """tests/basic_qp2/test_basic_qp2.cpp:10-70: eps 1e-6, polish_refine_iter 4;
    polish status NOT_PERFORMED (0) without polishing, SUCCESS (1) with it; x
    and y within TESTS_TOL of the stored solution."""
    p = pr.basic_qp2()
    s, r = gpu_solve(p, precision, eps_abs=1e-6, eps_rel=1e-6, polishing=polishing, polish_refine_iter=4)
    assert r.info.status_val == 1
    assert r.info.status_polish == expected
    assert np.max(np.abs(r.x - p["x"])) / np.max(np.abs(p["x"])) < 1e-4
    assert np.max(np.abs(r.y - p["y"])) / np.max(np.abs(p["y"])) < 1e-4
    assert abs(r.info.obj_val - p["obj"]) / abs(p["obj"]) < 1e-4


@pytest.mark.parametrize("precision", PRECISIONS)
def test_basic_qp_polish(precision):
    """tests/basic_qp/test_basic_qp.cpp:60-100: polishing on, scaling off, no
    warm start; x, y, objective within TESTS_TOL; the polish reaches the exact
    vertex (residuals below the ADMM ones)."""
    p = pr.basic_qp()
    s, r = gpu_solve(p, precision, polishing=1, scaling=0, warm_starting=0, max_iter=2000)
    assert r.info.status_val == 1
    assert r.info.status_polish == 1
    assert np.max(np.abs(r.x - p["x"])) < 1e-4
    assert np.max(np.abs(r.y - p["y"])) < 1e-4
    assert abs(r.info.obj_val - p["obj"]) < 1e-4


@pytest.mark.parametrize("precision", PRECISIONS)
def test_polish_mpc(precision):
    """Polishing on a problem with many active bounds: the polished point meets
    the eps test at least as well as the ADMM point (polish keeps the better)."""
    p = pr.mpc(15, seed=3)
    st = dict(eps_abs=1e-4, eps_rel=1e-4, max_iter=4000, scaling=10)
    s0, r0 = gpu_solve(p, precision, polishing=0, **st)
    s1, r1 = gpu_solve(p, precision, polishing=1, **st)
    assert r0.info.status_val == r1.info.status_val == 1
    assert r1.info.status_polish in (1, -1)
    c = pr.kkt_certificate(p, r1.x, r1.y)
    assert c["prim"] <= 10 * (1e-4 + 1e-4 * c["scale_p"])
    assert c["dual"] <= 10 * (1e-4 + 1e-4 * c["scale_d"])
    if r1.info.status_polish == 1:
        assert r1.info.prim_res <= r0.info.prim_res + 1e-12 or r1.info.dual_res <= r0.info.dual_res + 1e-12


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("name", ["basic_lp", "no_active_set", "unconstrained"])
def test_closed_forms(name, precision):
    p = getattr(pr, name)()
    s, r = gpu_solve(p, precision)
    T = tol_for(precision)
    assert r.info.status_val == 1, r.info.status
    assert np.max(np.abs(r.x - p["x"])) < T
    if len(p.get("y", [])):
        assert np.max(np.abs(r.y - p["y"])) < T
    assert abs(r.info.obj_val - p["obj"]) < T


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("case", [1, 2, 3, 4])
def test_primal_dual_infeasibility(case, precision):
    p = pr.primal_dual_infeasibility(case)
    s, r = gpu_solve(p, precision)
    assert r.info.status == p["status"]
    if case == 1:
        assert np.max(np.abs(r.x - p["x"])) < tol_for(precision)
    else:
        assert np.all(r.x == OSQP_NAN)
        cert = r.prim_inf_cert if "primal" in p["status"] else r.dual_inf_cert
        assert abs(np.max(np.abs(cert)) - 1.0) < 1e-12
        if "primal" in p["status"]:
            # certificate: A' y = 0 and u'y+ + l'y- < 0 (auxil.c:460-512)
            y = cert
            assert np.max(np.abs(p["A"].T @ y)) < 1e-3
            lc, uc = np.clip(p["l"], -1e30, 1e30), np.clip(p["u"], -1e30, 1e30)
            assert uc @ np.maximum(y, 0) + lc @ np.minimum(y, 0) < 0


@pytest.mark.parametrize("precision", PRECISIONS)
def test_primal_infeasibility_random(precision):
    p = pr.primal_infeasibility()
    s, r = gpu_solve(p, precision)
    assert r.info.status == "primal infeasible"


@pytest.mark.parametrize("precision", PRECISIONS)
def test_non_cvx_solve(precision):
    p = pr.non_cvx()
    s, r = gpu_solve(p, precision, adaptive_rho=0, sigma=5.0)
    assert r.info.status_val == 9
    assert r.info.obj_val == OSQP_NAN


@pytest.mark.parametrize("precision", PRECISIONS)
def test_large_qp_fixture(precision):
    p = pr.load_fixture(os.path.join(GOLDEN, "large_qp.json"))
    s, r = gpu_solve(p, precision)
    assert r.info.status_val == 1
    assert abs(r.info.obj_val - p["obj"]) / abs(p["obj"]) < tol_for(precision)


REF_SETS = ["hs21", "hs35", "qptest", "cvxqp2_s", "qpcblend", "largeqp", "primalc1"]
E4 = dict(eps_abs=1e-4, eps_rel=1e-4, max_iter=4000, scaling=10)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("name", REF_SETS)
def test_reference_problem_sets(name, precision):
    """examples/problems/*.c data sets, eps 1e-4."""
    p = pr.load_fixture(os.path.join(GOLDEN, f"{name}.json"))
    s, r = gpu_solve(p, precision, **E4)
    check_against_oracle(p, r, dict(FIXTURE, **E4), precision, name)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_qp_vs_oracle(seed, precision):
    """Random QPs of the random-QP configuration's shape (m = 10 n), eps 1e-4."""
    p = pr.random_qp(120, 1200, 0.15, 0.15, seed)
    s, r = gpu_solve(p, precision, **E4)
    check_against_oracle(p, r, dict(FIXTURE, **E4), precision, f"random_qp{seed}")
    c = pr.kkt_certificate(p, r.x, r.y)
    # the returned point passes the eps test OSQP used to stop (10x margin for
    # the FP64 recomputation here vs. the scaled/unscaled mix inside OSQP)
    assert c["prim"] <= 10 * (1e-4 + 1e-4 * c["scale_p"])
    assert c["dual"] <= 10 * (1e-4 + 1e-4 * c["scale_d"])


# Fixtures on which the FP64 engine and the oracle's PCG path take the same
# ADMM trajectory (same iterations, rho updates and point to rounding), and
# those on which it parts: on qpcblend and the 500-asset portfolio the CG's
# stopping iteration changes somewhere between two solves (different
# summation order of the CG dot products: a per-block tree on the GPU, a
# sequential sum in the oracle), and the two ill-conditioned trajectories
# then diverge; on largeqp and primalc1 they part late, with the same
# iteration count; on the lasso with n = 1000 (hybrid operators) a rho
# update falls on the other side of its threshold (the operators themselves
# are checked against scipy in test_gpu_operators.py).
TRAJ_CASES = {
    "basic_qp": (lambda: pr.basic_qp(), FIXTURE), "basic_qp2": (lambda: pr.basic_qp2(), FIXTURE),
    **{n: ((lambda n=n: pr.load_fixture(os.path.join(GOLDEN, f"{n}.json"))), dict(FIXTURE, **E4))
       for n in ["hs21", "hs35", "qptest", "cvxqp2_s", "primalc1"]},
    **{f"random_qp{k}": ((lambda k=k: pr.random_qp(120, 1200, 0.15, 0.15, k)), dict(FIXTURE, **E4))
       for k in range(3)},
    "lasso40": (lambda: pr.lasso(40, 400, 0.1, seed=1), dict(FIXTURE, **E4)),
    "lasso200": (lambda: pr.lasso(200, 2000, 0.02, seed=1), dict(FIXTURE, **E4)),
    "mpc10": (lambda: pr.mpc(10, seed=10), dict(FIXTURE, **E4)),
    "mpc30": (lambda: pr.mpc(30, seed=30), dict(FIXTURE, **E4)),
    "random_qp500": (lambda: pr.random_qp(500, 5000, 0.15, 0.15, seed=0), dict(eps_abs=1e-4, eps_rel=1e-4)),
    "largeqp": (lambda: pr.load_fixture(os.path.join(GOLDEN, "largeqp.json")), dict(FIXTURE, **E4)),
}
# same iteration count, points part late: largeqp, and primalc1, which runs to
# the 4000-iteration limit
TRAJ_PARTS = {"largeqp", "primalc1"}


@pytest.mark.parametrize("name", sorted(TRAJ_CASES))
def test_fp64_trajectory_matches_oracle_pcg(name):
    """The FP64 engine against the oracle's PCG path (the same algorithm in the
    same arithmetic): equal ADMM iterations, rho updates and status; the
    objective to 1e-9 and x, y to 1e-5 (relative) where the points agree."""
    make, st = TRAJ_CASES[name]
    p = make()
    s, r = gpu_solve(p, 64, **st)
    pc = snapshot(solve(p, 2, **dict(FIXTURE, **st)))
    e_obj, e_x, e_y = rel(r.info.obj_val, pc["obj"]), vec_err(r.x, pc["x"]), vec_err(r.y, pc["y"])
    print(f"{name}: it {r.info.iter}/{pc['iter']} rho updates {r.info.rho_updates}/{pc['rho_updates']} "
          f"obj {e_obj:.1e} x {e_x:.1e} y {e_y:.1e}")
    assert r.info.status_val == pc["status"]
    assert r.info.iter == pc["iter"] and r.info.rho_updates == pc["rho_updates"]
    if name not in TRAJ_PARTS:
        assert e_obj <= 1e-9 and e_x <= 1e-5 and e_y <= 1e-5


@pytest.mark.parametrize("precision", PRECISIONS)
def test_random_qp_config_full_size(precision):
    """The random-QP configuration at full size (n = 500, m = 5000, density
    0.15, P = M M' + 1e-2 I; default settings, eps 1e-4) against the oracle
    (check_against_oracle), and the GPU point passes the FP64 KKT certificate."""
    p = pr.random_qp(500, 5000, 0.15, 0.15, seed=0)
    st = dict(eps_abs=1e-4, eps_rel=1e-4)
    s, r = gpu_solve(p, precision, rho=0.1, alpha=1.6, max_iter=4000, scaling=10, **st)
    check_against_oracle(p, r, st, precision, "random_qp500")
    c = pr.kkt_certificate(p, r.x, r.y)
    assert c["prim"] <= 10 * (1e-4 + 1e-4 * c["scale_p"])
    assert c["dual"] <= 10 * (1e-4 + 1e-4 * c["scale_d"])
    s.close()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_multi_panel_qp(precision):
    """A QP wide enough for the column-panelled SpMV to use several panels in
    both precisions: A is 30000 x 9000 with 30 nonzeros per row, K = [P A']
    spans 39000 columns (5 FP32 panels of 8192, 10 FP64 panels of 4096; A
    takes 2 and 3).  Checked against the oracle's PCG path (same algorithm,
    FP64: objective to 1e-9 / 1e-6, x to 1e-6 / 1e-3 relative for the FP64 /
    FP32-PCG engine; the oracle's direct path would factor a dense 9000 x
    9000 Schur block here) and by the FP64 KKT certificate of the point."""
    p = pr.random_qp(9000, 30000, 2.0 / 9000, 30.0 / 9000, seed=7)
    st = dict(eps_abs=1e-4, eps_rel=1e-4, max_iter=4000, scaling=10)
    o = OracleSolver()
    fx = dict(FIXTURE)
    fx.update(st)
    assert o.setup(p["P"], p["q"], p["A"], p["l"], p["u"], linsys_solver=2, **fx) == 0
    o.solve()
    s, r = gpu_solve(p, precision, **st)
    print(f"multi-panel p{precision}: it {r.info.iter}/{o.info.iter} obj {rel(r.info.obj_val, o.info.obj_val):.1e} "
          f"x {vec_err(r.x, o.x):.1e} y {vec_err(r.y, o.y):.1e}")
    assert r.info.status_val == o.info.status_val == 1
    assert rel(r.info.obj_val, o.info.obj_val) <= (1e-9 if precision == 64 else 1e-6)
    assert vec_err(r.x, o.x) <= (1e-6 if precision == 64 else 1e-3)
    c = pr.kkt_certificate(p, r.x, r.y)
    assert c["prim"] <= 10 * (1e-4 + 1e-4 * c["scale_p"])
    assert c["dual"] <= 10 * (1e-4 + 1e-4 * c["scale_d"])


def test_large_qp_unspeculated_cg():
    """A QP whose CG iteration moves more than 64 MB (6e6 nonzeros in A), where
    the host queues no CG iteration before the previous one has decided to go
    on: both engines solve it, the FP32 engine's objective matches the FP64
    one to 1e-6, both points meet the FP64 KKT certificate, and a warm
    re-solve after a q update agrees again."""
    p = pr.random_qp(20000, 200000, 1e-4, 1.5e-3, seed=5)
    assert p["A"].nnz > 5_000_000
    st = dict(eps_abs=1e-4, eps_rel=1e-4, max_iter=4000)
    out = {}
    for prec in (64, 32):
        s, r = gpu_solve(p, prec, **st)
        assert r.info.status_val == 1, r.info.status
        c = pr.kkt_certificate(p, r.x, r.y)
        assert c["prim"] <= 10 * (1e-4 + 1e-4 * c["scale_p"])
        assert c["dual"] <= 10 * (1e-4 + 1e-4 * c["scale_d"])
        st_ = s.stats()
        assert st_.pcg_iters > 0 and st_.linsys_solves >= r.info.iter
        q2 = p["q"] + 0.1 * np.random.default_rng(1).standard_normal(p["q"].shape[0])
        s.update(q=q2)
        r2 = s.solve()
        assert r2.info.status_val == 1
        out[prec] = (r, r2)
        s.close()
    for k in (0, 1):
        a, b = out[32][k], out[64][k]
        print(f"solve {k}: FP32-PCG vs FP64 objective {rel(a.info.obj_val, b.info.obj_val):.1e}")
        assert rel(a.info.obj_val, b.info.obj_val) <= 1e-6


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("n,m_d,dens", [(40, 400, 0.1), (200, 2000, 0.02), (1000, 10000, 0.01)])
def test_lasso_vs_oracle(n, m_d, dens, precision):
    """The lasso configuration's QP (variables x, y, t; P = blkdiag(0, 2I, 0),
    so P has rows without a diagonal) at reduced size, against the oracle
    (check_against_oracle), and the KKT certificate.  The largest one runs
    its CG operators as hybrids (column panels over D's part, remainder CSR)."""
    p = pr.lasso(n, m_d, dens, seed=1)
    s, r = gpu_solve(p, precision, **E4)
    check_against_oracle(p, r, dict(FIXTURE, **E4), precision, f"lasso{n}")
    c = pr.kkt_certificate(p, r.x, r.y)
    assert c["prim"] <= 10 * (1e-4 + 1e-4 * c["scale_p"])
    assert c["dual"] <= 10 * (1e-4 + 1e-4 * c["scale_d"])


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("N", [10, 30])
def test_mpc_vs_oracle(N, precision):
    p = pr.mpc(N, seed=N)
    s, r = gpu_solve(p, precision, **E4)
    check_against_oracle(p, r, dict(FIXTURE, **E4), precision, f"mpc{N}")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_portfolio_update_q_warm(precision):
    """A 500-asset portfolio, then three new mu (update of q, warm re-solve),
    each against the oracle after the same update (check_against_oracle
    solves the updated problem from scratch for the optimum and the status;
    the warm sequences of both oracle paths give the same statuses)."""
    p = pr.portfolio(500, seed=3)
    st = dict(FIXTURE, **E4)
    s, r = gpu_solve(p, precision, **E4)
    check_against_oracle(p, r, st, precision, "portfolio500")
    o = solve(p, 1, **st)
    rng = np.random.default_rng(9)
    for j in range(3):
        q = pr.portfolio_q(rng.standard_normal(500), p["k"])
        s.update(q=q)
        r = s.solve()
        o.update_data_vec(q=q)
        o.solve()
        assert r.info.status_val == o.info.status_val == 1
        op = optimum(dict(p, q=q), **st)
        pc = solve(dict(p, q=q), 2, **st)
        e_gpu, e_pcg = rel(r.info.obj_val, op["obj"]), rel(pc.info.obj_val, op["obj"])
        print(f"portfolio500 p{precision} update {j}: obj vs optimum {e_gpu:.1e} (oracle PCG cold {e_pcg:.1e})")
        assert e_gpu <= e_pcg + OBJ_TOL


@pytest.mark.parametrize("precision", PRECISIONS)
def test_termination_fixed_iterations(precision):
    p = pr.basic_qp()
    s, r = gpu_solve(p, precision, max_iter=400, scaling=0, check_termination=0, warm_starting=0)
    assert r.info.iter == 400
    assert r.info.status_val == 1
    assert np.max(np.abs(r.x - p["x"])) < tol_for(precision)


def test_update_rho_same_iterations():
    p = pr.basic_qp()
    kw = dict(adaptive_rho=0, eps_abs=5e-5, eps_rel=5e-5, check_termination=1)
    a, ra = gpu_solve(p, 64, rho=0.7, **kw)
    b = OSQP(precision=64)
    b.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **dict(FIXTURE, rho=0.1, **kw))
    assert b.update_rho(0.7) == 0
    rb = b.solve()
    assert ra.info.status_val == rb.info.status_val == 1
    assert ra.info.iter == rb.info.iter


def test_warm_start_from_solution():
    p = pr.basic_qp()
    s, r = gpu_solve(p, 64, check_termination=1, adaptive_rho=0)
    s.warm_start(x=p["x"])
    s.warm_start(y=p["y"])
    r = s.solve()
    assert r.info.iter == 1
    assert r.info.status_val == 1


def test_time_limit():
    p = pr.basic_qp()
    s, r = gpu_solve(p, 64, rho=20, adaptive_rho=0)
    assert r.info.status_val == 1
    s.update_settings(time_limit=1e-7, eps_abs=1e-12, eps_rel=1e-12, max_iter=2_000_000_000, check_termination=0)
    s.cold_start()
    r = s.solve()
    assert r.info.status_val == 8


def test_settings_validation():
    """src/auxil.c:1040-1223: invalid settings are rejected at setup."""
    p = pr.basic_qp()
    for bad in (dict(alpha=2.0), dict(rho=0.0), dict(sigma=-1.0), dict(max_iter=0), dict(eps_abs=-1.0),
                dict(eps_abs=0.0, eps_rel=0.0), dict(cg_tol_fraction=1.0), dict(linsys_solver=1),
                dict(scaled_termination=2), dict(time_limit=0.0)):
        s = OSQP()
        assert s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **dict(FIXTURE, **bad)) == 2, bad
    s = OSQP()
    bad_l = p["u"] + 1.0
    assert s.setup(p["P"], p["q"], p["A"], bad_l, p["u"], **FIXTURE) == 1


def test_update_matrices():
    """tests/update_matrices: full P/A value update equals a fresh setup."""
    p = pr.random_qp(40, 200, 0.2, 0.2, 7)
    q = pr.random_qp(40, 200, 0.2, 0.2, 8)
    # same sparsity pattern, new values
    rng = np.random.default_rng(1)
    P2 = p["P"].copy()
    P2.data = P2.data * (1 + 0.1 * rng.random(P2.nnz))
    A2 = p["A"].copy()
    A2.data = rng.standard_normal(A2.nnz)
    st = dict(eps_abs=1e-5, eps_rel=1e-5)
    s, r = gpu_solve(p, 64, **st)
    assert s.update(Px=P2.data, Ax=A2.data) == 0
    r = s.solve()
    f, rf = gpu_solve(dict(p, P=P2, A=A2), 64, **st)
    assert r.info.status_val == rf.info.status_val == 1
    assert abs(r.info.obj_val - rf.info.obj_val) / abs(rf.info.obj_val) < 1e-4
    del q


@pytest.mark.parametrize("precision", PRECISIONS)
def test_long_rows_qp(precision):
    """A QP whose constraint matrix has rows longer than one SpMV task (a
    dense budget row over 6000 variables and 60 half-dense factor rows, as in
    the portfolio model).  The PCG path needs about 4100 ADMM iterations here
    (the oracle's restatement of the reference PCG, linsys_solver=2, FP64),
    so max_iter is 10000: the status equals the oracle PCG path's and direct
    path's, the objective is held to the reference's accuracy as in
    check_against_oracle, and the point meets the eps test on the primal and
    dual residuals."""
    p = pr.portfolio(6000, seed=4)
    st = dict(FIXTURE, **E4)
    st["max_iter"] = 10000
    pc = solve(p, 2, **st)
    di = solve(p, 1, **st)
    op = optimum(p, **st)
    s, r = gpu_solve(p, precision, **dict(E4, max_iter=10000))
    e_gpu, e_pcg = rel(r.info.obj_val, op["obj"]), rel(pc.info.obj_val, op["obj"])
    gb = gap_bound(p, op["x"], op["y"], 1e-4, 1e-4)
    print(f"long rows p{precision}: {r.info.status} it {r.info.iter} (oracle PCG {pc.info.status} it {pc.info.iter}) "
          f"obj vs optimum {e_gpu:.1e} (oracle PCG {e_pcg:.1e}, gap bound {gb:.1e})")
    assert r.info.status_val == pc.info.status_val == di.info.status_val == 1
    assert e_gpu <= max(gb, e_pcg) + OBJ_TOL
    c = pr.kkt_certificate(p, r.x, r.y)
    assert c["prim"] <= 10 * (1e-4 + 1e-4 * c["scale_p"])
    assert c["dual"] <= 10 * (1e-4 + 1e-4 * c["scale_d"])
