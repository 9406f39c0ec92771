"""Oracle answers shared by the GPU parity tests.

Three references, all from oracle/osqp_oracle.c in FP64:
  * direct(p): the reference's default path (builtin algebra, LDL' of the
    quasi-definite KKT matrix) at the test's own settings -- statuses are
    compared with it;
  * pcg(p): the restatement of the CUDA backend's PCG (linsys_solver = 2), the
    same algorithm the HIP engines run -- iterates, iteration counts and
    objectives agree with it to rounding while the trajectories agree;
  * optimum(p): the direct path at eps_abs = eps_rel = 1e-9, the FP64
    optimum against which objectives, x and y are measured, as the
    reference's tests measure theirs against held exact solutions
    (tests/osqp_tester.h:12-16, TESTS_TOL = 1e-4).

At eps 1e-4 the reference's own direct path can land up to ~2e-4 (relative)
away from the optimum, and its PCG path as far in another direction.  So a
result is held to OSQP's own accuracy promise: the termination test stops
when |gap| <= eps_abs + eps_rel * max(|x'Px|, |q'x|, |SC(y)|)
(src/auxil.c:231-306, 880-893), which bounds |f(x) - f*| by twice that
(gap_bound below); the oracle's direct path is asserted to meet the same
bound, so the bound is the reference's accuracy, not a looser one of ours.
"""
import numpy as np

from oracle import OracleSolver


def solve(p, linsys, q=None, **st):
    o = OracleSolver()
    assert o.setup(p["P"], p["q"] if q is None else q, p["A"], p["l"], p["u"], linsys_solver=linsys, **st) == 0
    o.solve()
    return o


def snapshot(o):
    i = o.info
    return dict(status=i.status_val, iter=i.iter, rho_updates=i.rho_updates, obj=i.obj_val,
                x=np.array(o.x, copy=True), y=np.array(o.y, copy=True))


def optimum(p, q=None, **st):
    """The direct path's solution at eps 1e-9 (1e-8 where that does not
    converge in 5 10^4 iterations); None when neither converges."""
    st = dict(st)
    st.pop("time_limit", None)
    for eps in (1e-9, 1e-8):
        st.update(eps_abs=eps, eps_rel=eps, max_iter=50000)
        o = solve(p, 1, q, **st)
        if o.info.status_val == 1:
            return snapshot(o)
    return None


def rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


def vec_err(a, b):
    """max |a - b| / max(1, max |b|)."""
    if len(b) == 0:
        return 0.0
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(1.0, float(np.max(np.abs(b)))))


def gap_bound(p, x_opt, y_opt, eps_abs, eps_rel, q=None):
    """Relative objective bound implied by OSQP's duality-gap stopping test
    at the optimum's scale: 2 (eps_abs + eps_rel max(|x'Px|, |q'x|, |SC|)) /
    max(1, |f*|), SC(y) = u'max(y, 0) + l'min(y, 0) over finite bounds."""
    q = p["q"] if q is None else q
    P = p["P"]
    Px = P @ x_opt + P.T @ x_opt - P.diagonal() * x_opt  # P holds the upper triangle
    xPx = float(x_opt @ Px)
    qx = float(q @ x_opt)
    l, u = np.asarray(p["l"]), np.asarray(p["u"])
    yp, yn = np.maximum(y_opt, 0), np.minimum(y_opt, 0)
    sc = float(np.sum(np.where(u < 1e20, u, 0) * yp) + np.sum(np.where(l > -1e20, l, 0) * yn))
    f = 0.5 * xPx + qx
    return 2 * (eps_abs + eps_rel * max(abs(xPx), abs(qx), abs(sc))) / max(1.0, abs(f))
