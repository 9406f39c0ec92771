"""Synthetic QP workloads (seeded) and the reference's closed-form test problems.

Workloads (shapes of the benchmark configurations):
  random_qp   n, m, P = M M' + 1e-2 I (M n x n sparse N(0,1)), A sparse N(0,1),
              l = -U(0,1), u = U(0,1), q ~ N(0,1)
  lasso       D m_d x n sparse N(0,1); variables (x, y, t); y = Dx - b;
              objective y'y + lambda 1't, -t <= x <= t   (docs/examples/lasso.rst)
  mpc         nx=12, nu=4, horizon N; sparse formulation (docs/examples/mpc.rst)
  portfolio   n assets, k = n/100 factors (docs/examples/portfolio.rst)

Closed-form problems with known solutions are those the reference's tests
generate (tests/*/generate_problem.py); their expected values are quoted there.
"""
import numpy as np
import scipy.sparse as sp


def sprandn(m, n, density, rng):
    """m x n sparse matrix with ~density*m*n N(0,1) entries, CSC, sorted, no duplicates.

    Draws a Binomial(m, density) count per column and that many distinct rows;
    for large matrices sampling is done with replacement and duplicates are
    dropped (the expected loss at density 2e-3 is 0.1 %).
    """
    counts = rng.binomial(m, density, size=n)
    total = int(counts.sum())
    if total == 0:
        return sp.csc_matrix((m, n))
    if m * n <= 5e7:
        M = sp.random(m, n, density=density, format="csc", random_state=rng, data_rvs=rng.standard_normal)
        M.sort_indices()
        return M
    cols = np.repeat(np.arange(n, dtype=np.int64), counts)
    rows = rng.integers(0, m, size=total, dtype=np.int64)
    key = np.unique(cols * m + rows)
    cols = (key // m).astype(np.int32)
    rows = (key % m).astype(np.int32)
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, cols.astype(np.int64) + 1, 1)
    indptr = np.cumsum(indptr)
    data = rng.standard_normal(len(rows))
    return sp.csc_matrix((data, rows, indptr.astype(np.int32)), shape=(m, n))


def random_qp(n, m, dens_M, dens_A, seed=0):
    """Random sparse QP of the random-QP configurations (P = MM' + 1e-2 I)."""
    rng = np.random.default_rng(seed)
    M = sprandn(n, n, dens_M, rng)
    P = (M @ M.T + 1e-2 * sp.eye(n)).tocsc()
    P = sp.triu(P, format="csc")
    q = rng.standard_normal(n)
    A = sprandn(m, n, dens_A, rng)
    l = -rng.random(m)
    u = rng.random(m)
    return dict(P=P, q=q, A=A, l=l, u=u, name=f"random_qp_n{n}_m{m}_s{seed}")


def lasso(n, m_d, density, seed=0):
    """Lasso as a QP: minimise y'y + lam 1't  s.t.  y = Dx - b, -t <= x <= t."""
    rng = np.random.default_rng(seed)
    D = sprandn(m_d, n, density, rng)
    x_true = np.where(rng.random(n) < 0.1, rng.standard_normal(n) / np.sqrt(n), 0.0)
    b = D @ x_true + 0.1 * rng.standard_normal(m_d)
    lam = 0.2 * np.linalg.norm(D.T @ b, np.inf)
    In, Im = sp.eye(n, format="csc"), sp.eye(m_d, format="csc")
    P = sp.block_diag([sp.csc_matrix((n, n)), 2 * Im, sp.csc_matrix((n, n))], format="csc")
    q = np.hstack([np.zeros(n + m_d), lam * np.ones(n)])
    A = sp.bmat([[D, -Im, None], [In, None, -In], [In, None, In]], format="csc")
    l = np.hstack([b, -np.inf * np.ones(n), np.zeros(n)])
    u = np.hstack([b, np.zeros(n), np.inf * np.ones(n)])
    return dict(P=sp.triu(P, format="csc"), q=q, A=A, l=l, u=u, name=f"lasso_n{n}_m{m_d}_s{seed}")


def mpc(N, seed=0, nx=12, nu=4):
    """Linear MPC, sparse formulation (docs/examples/mpc.rst:70-87).

    A_d ~ N(0,1) rescaled to spectral radius 0.95, B_d ~ N(0,1), Q = QN = I,
    R = 0.1 I, |u| <= 1, |x| <= 5, x0 ~ U(-1,1), reference state 0.

    The sparsity pattern depends on N only, so it is built once per (N, nx,
    nu) by mpc_build with marker values in place of A_d and B_d, and each
    request fills its values through the markers' positions: the same
    matrices, entry for entry, as mpc_build gives (about 50x faster than
    building them with scipy each time; a stream benchmark makes 10^4+ of them).
    """
    rng = np.random.default_rng(seed)
    Ad = rng.standard_normal((nx, nx))
    Ad *= 0.95 / max(abs(np.linalg.eigvals(Ad)))
    Bd = rng.standard_normal((nx, nu))
    x0 = rng.uniform(-1, 1, nx)
    t = _mpc_template(N, nx, nu)
    vals = np.concatenate([t["fixed"], Ad.ravel(), Bd.ravel()])
    A = sp.csc_matrix((vals[t["src"]], t["indices"], t["indptr"]), shape=t["shape"])
    l = t["l"].copy()
    u = t["u"].copy()
    l[:nx] = -x0
    u[:nx] = -x0
    return dict(P=t["P"].copy(), q=np.zeros(t["P"].shape[0]), A=A, l=l, u=u, name=f"mpc_N{N}_s{seed}")


_MPC_TEMPLATES = {}


def _mpc_template(N, nx, nu):
    key = (N, nx, nu)
    t = _MPC_TEMPLATES.get(key)
    if t is not None:
        return t
    # markers: A_d[i, j] -> 10 + i*nx + j, B_d[i, j] -> 10 + nx*nx + i*nu + j;
    # the fixed entries (+-1 of the identities) keep their values, mapped to
    # slots 0 (-1) and 1 (+1) of the value table
    Am = 10.0 + np.arange(nx * nx, dtype=float).reshape(nx, nx)
    Bm = 10.0 + nx * nx + np.arange(nx * nu, dtype=float).reshape(nx, nu)
    ref = mpc_build(N, Am, Bm, np.zeros(nx), nx, nu)
    A = ref["A"]
    d = A.data
    src = np.where(d == -1.0, 0, np.where(d == 1.0, 1, d.astype(np.int64) - 10 + 2))
    assert np.all((d == -1.0) | (d == 1.0) | ((d >= 10) & (d < 10 + nx * nx + nx * nu)))
    t = dict(fixed=np.array([-1.0, 1.0]), src=src, indices=A.indices.copy(), indptr=A.indptr.copy(),
             shape=A.shape, P=ref["P"], l=ref["l"], u=ref["u"])
    _MPC_TEMPLATES[key] = t
    return t


def mpc_build(N, Ad, Bd, x0, nx=12, nu=4):
    """The MPC QP for given A_d, B_d, x0, built with scipy (docs/examples/mpc.rst:70-87)."""
    Q = sp.eye(nx)
    QN = sp.eye(nx)
    R = 0.1 * sp.eye(nu)
    P = sp.block_diag([sp.kron(sp.eye(N), Q), QN, sp.kron(sp.eye(N), R)], format="csc")
    q = np.zeros((N + 1) * nx + N * nu)
    Ax = sp.kron(sp.eye(N + 1), -sp.eye(nx)) + sp.kron(sp.eye(N + 1, k=-1), sp.csc_matrix(Ad))
    Bu = sp.kron(sp.vstack([sp.csc_matrix((1, N)), sp.eye(N)]), sp.csc_matrix(Bd))
    Aeq = sp.hstack([Ax, Bu])
    leq = np.hstack([-x0, np.zeros(N * nx)])
    Aineq = sp.eye((N + 1) * nx + N * nu)
    lineq = np.hstack([np.kron(np.ones(N + 1), -5 * np.ones(nx)), np.kron(np.ones(N), -np.ones(nu))])
    uineq = np.hstack([np.kron(np.ones(N + 1), 5 * np.ones(nx)), np.kron(np.ones(N), np.ones(nu))])
    A = sp.vstack([Aeq, Aineq], format="csc")
    A.eliminate_zeros()
    l = np.hstack([leq, lineq])
    u = np.hstack([leq, uineq])
    return dict(P=sp.triu(P, format="csc"), q=q, A=A, l=l, u=u)


def portfolio(n, seed=0, gamma=1.0):
    """Factor-model portfolio (docs/examples/portfolio.rst:45-63); returns mu separately."""
    rng = np.random.default_rng(seed)
    k = max(1, n // 100)
    F = sp.random(n, k, density=0.5, format="csc", random_state=rng, data_rvs=rng.standard_normal)
    Dd = sp.diags(rng.random(n) * np.sqrt(k), format="csc")
    mu = rng.standard_normal(n)
    P = sp.block_diag([Dd, sp.eye(k)], format="csc")
    q = np.hstack([-mu / (2 * gamma), np.zeros(k)])
    A = sp.bmat([[F.T, -sp.eye(k)], [sp.csc_matrix(np.ones((1, n))), None], [sp.eye(n), None]], format="csc")
    l = np.hstack([np.zeros(k), 1.0, np.zeros(n)])
    u = np.hstack([np.zeros(k), 1.0, np.ones(n)])
    return dict(P=sp.triu(P, format="csc"), q=q, A=A, l=l, u=u, mu=mu, k=k, gamma=gamma, name=f"portfolio_n{n}_s{seed}")


def portfolio_q(mu, k, gamma=1.0):
    return np.hstack([-mu / (2 * gamma), np.zeros(k)])


# ---- the reference's closed-form problems -----------------------------------------

def basic_qp():
    """tests/basic_qp/generate_problem.py:5-27"""
    return dict(P=sp.triu(sp.csc_matrix([[4., 1.], [1., 2.]]), format="csc"), q=np.ones(2),
                A=sp.csc_matrix(np.array([[1., 1.], [1., 0.], [0., 1.], [0., 1.]])),
                l=np.array([1., 0., 0., -np.inf]), u=np.array([1., 0.7, 0.7, np.inf]),
                x=np.array([0.3, 0.7]), y=np.array([-2.9, 0.0, 0.2, 0.0]), obj=1.88,
                q_new=np.array([2.5, 3.2]), l_new=np.array([0.8, -3.4, -np.inf, 0.5]),
                u_new=np.array([1.6, 1.0, np.inf, 0.5]))


def basic_qp2():
    """tests/basic_qp2/generate_problem.py:5-30"""
    return dict(P=sp.triu(sp.csc_matrix([[11., 0.], [0., 0.]]), format="csc"), q=np.array([3., 4.]),
                A=sp.csc_matrix(np.array([[-1., 0.], [0., -1.], [-1., 3.], [2., 5.], [3., 4]])),
                l=-np.inf * np.ones(5), u=np.array([0., 0., -15., 100., 80.]),
                x=np.array([15., -0.]), y=np.array([0., 508., 168., 0., 0.]), obj=1282.5,
                q_new=np.array([1., 1.]), u_new=np.array([-2., 0., -20., 100., 80.]),
                x_new=np.array([20., -0.]), y_new=np.array([0., 664., 221., 0., 0.]), obj_new=2220.0)


def basic_lp():
    """tests/basic_lp/generate_problem.py:5-27"""
    return dict(P=sp.csc_matrix((2, 2)), q=np.array([1.1, 0.95]),
                A=sp.csc_matrix(np.array([[1., 1.], [1., 0.], [0., 1.], [0., 1.]])),
                l=np.array([1., 0., 0., -np.inf]), u=np.array([1., 0.7, 0.7, np.inf]),
                x=np.array([0.3, 0.7]), y=np.array([-1.1, 0.0, 0.15, 0.0]), obj=0.995)


def unconstrained():
    """tests/unconstrained/generate_problem.py:5-14"""
    return dict(P=sp.diags([0.617022, 0.92032449, 0.20011437, 0.50233257, 0.34675589], format="csc"),
                q=np.array([-1.10593508, -1.65451545, -2.3634686, 1.13534535, -1.01701414]),
                A=sp.csc_matrix((0, 5)), l=np.zeros(0), u=np.zeros(0),
                x=np.array([1.79237542, 1.79775228, 11.81058885, -2.26014678, 2.93293975]),
                obj=-19.209752026813277)


def no_active_set():
    """tests/no_active_set/generate_problem.py:5-19"""
    return dict(P=sp.triu(sp.csc_matrix([[11., 0.], [0., 5.]]), format="csc"), q=np.zeros(2),
                A=sp.csc_matrix(np.array([[1., 1.], [1., 0.], [0., 1.]])),
                l=np.array([-1., -0.5, -0.25]), u=np.array([1., 0.7, 0.7]),
                x=np.zeros(2), y=np.zeros(3), obj=0.0)


def primal_dual_infeasibility(case):
    """tests/primal_dual_infeasibility/generate_problem.py:5-33 (cases 1..4)"""
    P = sp.diags([1., 0.], format="csc")
    q = np.array([1., -1.])
    A12 = sp.csc_matrix([[1., 1.], [1., 0.], [0., 1.]])
    A34 = sp.csc_matrix([[1., 0.], [1., 0.], [0., 1.]])
    l = np.array([0., 1., 1.])
    u = {1: np.array([5., 3., 3.]), 2: np.array([0., 3., 3.]), 3: np.array([2., 3., np.inf]),
         4: np.array([0., 3., np.inf])}[case]
    status = {1: "solved", 2: "primal infeasible", 3: "dual infeasible", 4: "primal infeasible"}[case]
    d = dict(P=P, q=q, A=A12 if case <= 2 else A34, l=l, u=u, status=status)
    if case == 1:
        d.update(x=np.array([1., 3.]), y=np.array([0., -2., 1.]), obj=-1.5)
    return d


def non_cvx():
    """tests/non_cvx/generate_problem.py:5-12"""
    return dict(P=sp.triu(sp.csc_matrix([[2., 5.], [5., 1.]]), format="csc"), q=np.array([3., 4.]),
                A=sp.csc_matrix([[-1., 0.], [0., -1.], [-1., 3.], [2., 5.], [3., 4]]),
                l=-np.inf * np.ones(5), u=np.array([0., 0., -15., 100., 80.]))


def primal_infeasibility(seed=2):
    """Construction of tests/primal_infeasibility/generate_problem.py:8-30 (own draws):
    row n/2 duplicates row n/2+1 with a disjoint interval -> primal infeasible."""
    rng = np.random.default_rng(seed)
    n, m = 50, 150
    Pt = sp.random(n, n, random_state=rng)
    P = sp.triu(Pt.T @ Pt + sp.eye(n), format="csc")
    q = rng.standard_normal(n)
    A = sp.random(m, n, random_state=rng).tolil()
    u = 3 + rng.standard_normal(m)
    l = -3 + rng.standard_normal(m)
    A[n // 2, :] = A[n // 2 + 1, :]
    l[n // 2] = u[n // 2 + 1] + 10 * rng.random()
    u[n // 2] = l[n // 2] + 0.5
    return dict(P=P, q=q, A=A.tocsc(), l=l, u=u, status="primal infeasible")


def update_matrices_problem():
    """The data of the reference's matrix-update test
    (tests/update_matrices/generate_problem.py:6-64, PCG64(1) draws through
    scipy.sparse.random in the same order), with the solutions it holds
    (:86-111) for the original problem and after new P, new A, and both."""
    from numpy.random import PCG64, Generator
    rg = Generator(PCG64(1))
    n, m, dens = 5, 8, 0.7
    A = sp.random(m, n, density=dens, format="csc", random_state=rg)
    P = sp.random(n, n, density=dens, random_state=rg)
    P = (P @ P.T).tocsc() + sp.eye(n, format="csc")
    Pu = sp.triu(P, format="csc")
    A_new = A.copy()
    A_new.data += rg.standard_normal(A_new.nnz)
    Pu_new = Pu.copy()
    Pu_new.data += 0.1 * rg.standard_normal(Pu_new.nnz)
    q = rg.standard_normal(n)
    l = -30 + rg.standard_normal(m)
    u = 30 + rg.standard_normal(m)
    x = np.array([-0.34282147, 0.0260615, 0.27197987, -0.01202531, 0.44052732])
    x_P = np.array([-0.36033252, 0.10729745, 0.22965285, 0.06323582, 0.42301398])
    return dict(P=Pu, q=q, A=A, l=l, u=u, Pu_new=Pu_new, A_new=A_new,
                x=x, obj=-0.37221, x_P_new=x_P, obj_P_new=-0.36765, x_A_new=x, obj_A_new=-0.37221,
                x_PA_new=x_P, obj_PA_new=-0.36765, y=np.zeros(m))


def load_fixture(path):
    """Reference data set extracted to JSON (tests/golden/*.json)."""
    import json
    with open(path) as f:
        fx = json.load(f)
    n, m = fx["n"], fx["m"]
    P = sp.csc_matrix((fx["P"]["x"], fx["P"]["i"], fx["P"]["p"]), shape=(n, n))
    A = sp.csc_matrix((fx["A"]["x"], fx["A"]["i"], fx["A"]["p"]), shape=(m, n))
    return dict(P=P, q=np.array(fx["q"]), A=A, l=np.array(fx["l"]), u=np.array(fx["u"]),
                obj=fx.get("obj_val"), name=fx["name"])


def kkt_certificate(prob, x, y):
    """FP64 KKT residuals of a returned (x, y) on the unscaled problem:
    primal ||Ax - proj(Ax)||_inf, dual ||Px + q + A'y||_inf, objective."""
    P = prob["P"]
    Pf = (P + sp.triu(P, 1).T).tocsc()
    A = prob["A"]
    l = np.clip(prob["l"], -1e30, 1e30)
    u = np.clip(prob["u"], -1e30, 1e30)
    Ax = A @ x
    z = np.minimum(np.maximum(Ax, l), u)
    prim = np.max(np.abs(Ax - z)) if len(z) else 0.0
    Px = Pf @ x
    dual = np.max(np.abs(Px + prob["q"] + A.T @ y))
    obj = 0.5 * x @ Px + prob["q"] @ x
    scale_p = max(np.max(np.abs(Ax)) if len(Ax) else 0.0, np.max(np.abs(z)) if len(z) else 0.0)
    scale_d = max(np.max(np.abs(Px)), np.max(np.abs(prob["q"])), np.max(np.abs(A.T @ y)) if A.shape[0] else 0.0)
    return dict(prim=prim, dual=dual, obj=obj, scale_p=scale_p, scale_d=scale_d)
