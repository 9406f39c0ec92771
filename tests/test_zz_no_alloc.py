"""osqp_amd_reduce_noalloc: the reductions that allocate nothing.

It runs the queued reductions' kernel with its partials and tickets in a
fixed device array, on the caller's stream, writing to caller-owned device
memory.  Checked against the allocating entry points of the same kinds
(OSQPVectorf_norm_inf, scaled_norm_inf, norm_inf_diff, norm_1, dot_prod,
dot_prod_signed): the same bits, on two calls in a row, the second with
smaller inputs and fewer reductions (the shared array keeps the first call's
contents), and at the 16-reduction maximum; requests beyond it are refused.
"""
import ctypes as C

import numpy as np
import pytest

from osqp_amd.algebra import Vec

pytestmark = pytest.mark.gpu

NORM_INF, SCALED_NORM_INF, NORM_INF_DIFF, NORM_1, DOT, DOT_POS, DOT_NEG = range(7)


class Item(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("a", C.c_void_p), ("b", C.c_void_p)]


def entry(L):
    f = L.osqp_amd_reduce_noalloc
    f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    L.OSQPVectorf_device_ptr.argtypes = [C.c_void_p]
    L.OSQPVectorf_device_ptr.restype = C.c_void_p
    return f


def hip_sync():
    hip = C.CDLL("libamdhip64.so.7")
    assert hip.hipDeviceSynchronize() == 0


def allocating(L, kind, va, vb):
    if kind == NORM_INF:
        return L.OSQPVectorf_norm_inf(va.h)
    if kind == SCALED_NORM_INF:
        return L.OSQPVectorf_scaled_norm_inf(va.h, vb.h)
    if kind == NORM_INF_DIFF:
        return L.OSQPVectorf_norm_inf_diff(va.h, vb.h)
    if kind == NORM_1:
        return L.OSQPVectorf_norm_1(va.h)
    if kind == DOT:
        return L.OSQPVectorf_dot_prod(va.h, vb.h)
    return L.OSQPVectorf_dot_prod_signed(va.h, vb.h, 1 if kind == DOT_POS else -1)


def run_and_check(algebra, f, L, specs, seed):
    """specs: list of (kind, n).  Runs them in one no-alloc call and compares
    each result with the allocating entry point's, bit for bit."""
    rng = np.random.default_rng(seed)
    vecs = [(Vec(rng.standard_normal(n)), Vec(rng.standard_normal(n))) for _, n in specs]
    items = (Item * len(specs))()
    for k, ((kind, n), (va, vb)) in enumerate(zip(specs, vecs)):
        items[k] = Item(kind, n, L.OSQPVectorf_device_ptr(va.h), L.OSQPVectorf_device_ptr(vb.h))
    out = Vec(np.full(len(specs), np.nan))
    L.osqp_amd_algebra_sync()
    print(f"no-alloc call: {len(specs)} reductions, n = {sorted(set(n for _, n in specs))}", flush=True)
    assert f(items, len(specs), L.OSQPVectorf_device_ptr(out.h), None) == 0
    hip_sync()
    got = out.get()
    for k, ((kind, n), (va, vb)) in enumerate(zip(specs, vecs)):
        want = allocating(L, kind, va, vb)
        assert got[k] == want, (k, kind, n, got[k], want)
    for va, vb in vecs:
        va.free()
        vb.free()
    out.free()


def test_no_alloc_reductions_match_allocating(algebra):
    L = algebra.lib()
    f = entry(L)
    # first call: every kind, lengths that take 1 to 512 blocks, and an empty vector
    big = [(k, 1_000_003) for k in range(7)] + [(k, 65_537) for k in range(7)] + [(DOT, 0), (NORM_INF, 5)]
    assert len(big) == 16
    run_and_check(algebra, f, L, big, seed=1)
    # second call, smaller: the tickets and partials left by the first are reused
    run_and_check(algebra, f, L, [(DOT, 777), (NORM_INF_DIFF, 3), (NORM_1, 1)], seed=2)


def test_no_alloc_refuses_beyond_maximum(algebra):
    L = algebra.lib()
    f = entry(L)
    v = Vec(np.ones(4))
    out = Vec(np.zeros(17))
    items = (Item * 17)(*[Item(NORM_INF, 4, L.OSQPVectorf_device_ptr(v.h), None) for _ in range(17)])
    assert f(items, 17, L.OSQPVectorf_device_ptr(out.h), None) != 0
    assert f(items, 0, L.OSQPVectorf_device_ptr(out.h), None) != 0
    bad = (Item * 1)(Item(99, 4, L.OSQPVectorf_device_ptr(v.h), None))
    assert f(bad, 1, L.OSQPVectorf_device_ptr(out.h), None) != 0
    assert np.all(out.get() == 0.0)  # nothing was launched
    v.free()
    out.free()
