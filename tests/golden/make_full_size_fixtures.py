"""Writes the oracle's answers for the two configurations too large to solve
on the CPU inside a test: the large random QP (n = 50000, m = 500000, A of
density 2e-3, P = M M' + 1e-2 I, seed 0) and the lasso QP (n = 20000,
m_d = 200000, D of density 1e-3, seed 0), both at eps_abs = eps_rel = 1e-4
with the library's default settings.

The oracle (oracle/osqp_oracle.c, FP64) runs its PCG path (linsys_solver = 2,
the restatement of algebra/cuda/lin_sys/indirect/cuda_pcg*.cu): the reference's
direct solver would factor a nearly dense Schur block of A'A here.  Stored per
configuration (tests/golden/<name>_oracle.npz):
  * data checksums (nnz, sums of P, A, q, l, u) so a test can tell that it
    regenerated the same problem;
  * status, iterations, rho updates, objective, residuals, total CG iterations;
  * x in full, and y in full (both compress well: y is zero on inactive rows).

    python tests/golden/make_full_size_fixtures.py [large_qp|lasso ...]

Takes about 10 minutes on one core for the large QP.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import OracleSolver  # noqa: E402
from osqp_amd import problems as pr  # noqa: E402

CONFIGS = {
    "large_qp": lambda: pr.random_qp(50000, 500000, 1e-4, 2e-3, seed=0),
    "lasso": lambda: pr.lasso(20000, 200000, 1e-3, seed=0),
}
SETTINGS = dict(eps_abs=1e-4, eps_rel=1e-4, verbose=0)


def checksums(p):
    return np.array([p["P"].nnz, p["A"].nnz, np.sum(p["P"].data), np.sum(p["A"].data), np.sum(p["q"]),
                     np.sum(np.clip(p["l"], -1e30, 1e30)), np.sum(np.clip(p["u"], -1e30, 1e30))])


def make(name):
    t0 = time.time()
    p = CONFIGS[name]()
    print(f"{name}: generated n={p['P'].shape[0]} m={p['A'].shape[0]} nnz(A)={p['A'].nnz} "
          f"in {time.time() - t0:.1f}s", flush=True)
    o = OracleSolver()
    assert o.setup(p["P"], p["q"], p["A"], p["l"], p["u"], linsys_solver=2, **SETTINGS) == 0
    t0 = time.time()
    o.solve()
    el = time.time() - t0
    inf = o.info
    print(f"{name}: status {inf.status_val} iter {inf.iter} rho_updates {inf.rho_updates} "
          f"obj {inf.obj_val:.12e} ({el:.0f}s, {o.total_pcg_iters()} CG iterations)", flush=True)
    np.savez_compressed(
        os.path.join(HERE, f"{name}_oracle.npz"), checksums=checksums(p),
        info=np.array([inf.status_val, inf.iter, inf.rho_updates, inf.obj_val, inf.prim_res, inf.dual_res,
                       inf.duality_gap, o.total_pcg_iters(), el]),
        x=o.x, y=o.y)


if __name__ == "__main__":
    for name in (sys.argv[1:] or list(CONFIGS)):
        make(name)
