"""The Python front end's marshalling helper (osqp_amd/_marshal.so), on the CPU.

csc_batch packs QPs into osqp_batch_add_many's arguments without copying:
the OSQPCscMatrix structs must describe the caller's own arrays, the pointer
tables must follow the argument order (P, q, A, l, u, x0, y0), and any problem
that is not in the engine's exact form must make it return None, so that
BatchEngine.add_many converts the arrays itself (as CscHolder does).
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from osqp_amd import _types as T
from osqp_amd import problems as pr

_marshal = pytest.importorskip("osqp_amd._marshal")


def _tables(out, k):
    buf, base = out
    ps, ts = C.sizeof(T.OSQPCscMatrix), C.sizeof(C.c_void_p)
    tab = base + 2 * k * ps
    cols = [(C.c_void_p * k).from_address(tab + c * k * ts) for c in range(7)]
    dims = np.frombuffer(buf, dtype=np.int32, count=2 * k, offset=2 * k * ps + 7 * k * ts)
    return cols, dims


def _addr(a):
    return None if a is None else a.ctypes.data


def test_layout_points_at_the_callers_arrays():
    probs = [pr.mpc(N, seed=N) for N in (10, 17, 30)] + [pr.portfolio(500, seed=3), pr.basic_qp()]
    probs[1] = dict(probs[1], x0=np.linspace(0, 1, probs[1]["P"].shape[0]))
    probs[2] = dict(probs[2], y0=np.ones(probs[2]["A"].shape[0]))
    probs = [dict(p, P=sp.csc_matrix(p["P"]), A=sp.csc_matrix(p["A"])) for p in probs]
    for p in probs:
        p["P"].sort_indices()
        p["A"].sort_indices()
        for key in ("q", "l", "u"):
            p[key] = np.ascontiguousarray(p[key], dtype=np.float64)
    out = _marshal.csc_batch(probs)
    assert out is not None
    k = len(probs)
    (Pt, qt, At, lt, ut, xt, yt), dims = _tables(out, k)
    for j, p in enumerate(probs):
        for tab, key in ((Pt, "P"), (At, "A")):
            s = T.OSQPCscMatrix.from_address(tab[j])
            M = p[key]
            assert (s.m, s.n, s.nzmax) == (M.shape[0], M.shape[1], M.nnz)
            assert C.cast(s.p, C.c_void_p).value == M.indptr.ctypes.data
            assert C.cast(s.i, C.c_void_p).value == M.indices.ctypes.data
            assert C.cast(s.x, C.c_void_p).value == M.data.ctypes.data
        assert qt[j] == _addr(p["q"]) and lt[j] == _addr(p["l"]) and ut[j] == _addr(p["u"])
        assert xt[j] == _addr(p.get("x0")) and yt[j] == _addr(p.get("y0"))
        assert dims[j] == p["P"].shape[0] and dims[k + j] == p["A"].shape[0]


def _base():
    p = pr.mpc(12, seed=4)
    return dict(p, P=sp.csc_matrix(p["P"]), A=sp.csc_matrix(p["A"]))


def _unsorted(M):
    M = M.copy()
    j = next(j for j in range(M.shape[1]) if M.indptr[j + 1] - M.indptr[j] >= 2)
    s = M.indptr[j]
    M.indices[s], M.indices[s + 1] = M.indices[s + 1], M.indices[s]
    return M


@pytest.mark.parametrize("case", ["unsorted A", "unsorted P", "csr A", "short q", "long l", "float32 u",
                                  "strided q", "short x0", "missing u", "dense A", "decreasing indptr"])
def test_rejects_what_the_engine_does_not_take(case):
    p = _base()
    n, m = p["P"].shape[0], p["A"].shape[0]
    if case == "unsorted A":
        p["A"] = _unsorted(p["A"])
    elif case == "unsorted P":
        P = sp.csc_matrix(sp.triu(sp.random(n, n, density=0.3, random_state=1) + sp.eye(n)))
        P.sort_indices()
        p["P"] = _unsorted(P)
    elif case == "csr A":
        p["A"] = p["A"].tocsr()
    elif case == "short q":
        p["q"] = p["q"][:-1]
    elif case == "long l":
        p["l"] = np.append(p["l"], 0.0)
    elif case == "float32 u":
        p["u"] = p["u"].astype(np.float32)
    elif case == "strided q":
        p["q"] = np.repeat(p["q"], 2)[::2]
    elif case == "short x0":
        p["x0"] = np.zeros(n - 1)
    elif case == "missing u":
        del p["u"]
    elif case == "dense A":
        p["A"] = p["A"].toarray()
    elif case == "decreasing indptr":
        A = p["A"].copy()
        A.indptr = A.indptr.copy()
        A.indptr[2], A.indptr[1] = A.indptr[1], A.indptr[2] + 1
        p["A"] = A
    assert _marshal.csc_batch([p]) is None
    # one bad problem sends the whole list down the converting path
    assert _marshal.csc_batch([_base(), p]) is None


def test_sortedness_matches_scipy():
    rng = np.random.default_rng(7)
    for t in range(200):
        M = sp.random(int(rng.integers(1, 30)), int(rng.integers(1, 30)), density=float(rng.uniform(0, 0.7)),
                      format="csc", random_state=t)
        if rng.integers(2):
            M.sort_indices()
        elif M.nnz:
            # shuffle within one column
            j = int(rng.integers(M.shape[1]))
            s, e = M.indptr[j], M.indptr[j + 1]
            M.indices[s:e] = rng.permutation(M.indices[s:e])
        n = M.shape[1]
        p = dict(P=sp.csc_matrix(sp.eye(n)), A=M, q=np.zeros(n), l=np.zeros(M.shape[0]), u=np.zeros(M.shape[0]))
        ok = all(np.all(np.diff(M.indices[M.indptr[j]:M.indptr[j + 1]]) >= 0) for j in range(n))
        assert (_marshal.csc_batch([p]) is not None) == ok
