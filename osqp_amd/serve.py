"""Python front end of the batched serving engine (include/osqp_amd_serve.h).

    eng = BatchEngine(precision=32, eps_abs=1e-4, eps_rel=1e-4)
    ids = [eng.add(p["P"], p["q"], p["A"], p["l"], p["u"]) for p in problems]
    results = {}
    while eng.pending():
        for i in eng.step():           # one epoch on the GPU (pipelined)
            results[i] = eng.poll(i)   # polling frees the solve's memory

Settings are OSQP's (keyword arguments as in OSQP.setup).  All work happens in
libosqp_amd.so; this module marshals arrays only.
"""
import ctypes as C
from types import SimpleNamespace

import numpy as np

from . import _types as T
from .solver import load_library

try:  # the marshalling helper built beside libosqp_amd.so (osqp_amd/build.py)
    from . import _marshal
except ImportError:  # arrays are then converted one by one in Python
    _marshal = None

_bound = False


class BatchInfo(C.Structure):
    _fields_ = [
        ("status_val", T.OSQPInt), ("iter", T.OSQPInt), ("rho_updates", T.OSQPInt),
        ("obj_val", T.OSQPFloat), ("dual_obj_val", T.OSQPFloat), ("prim_res", T.OSQPFloat),
        ("dual_res", T.OSQPFloat), ("duality_gap", T.OSQPFloat), ("rho_estimate", T.OSQPFloat),
        ("primdual_int", T.OSQPFloat), ("rel_kkt_error", T.OSQPFloat), ("latency", T.OSQPFloat),
        ("cg_iters", C.c_longlong),
    ]


def _bind(lib):
    global _bound
    if _bound:
        return lib
    vp = C.c_void_p
    lib.osqp_batch_create.argtypes = [C.POINTER(vp), vp, C.c_int, C.c_int, C.c_longlong]
    lib.osqp_batch_create.restype = C.c_int
    lib.osqp_batch_add.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.c_int, C.POINTER(C.c_int)]
    lib.osqp_batch_add.restype = C.c_int
    lib.osqp_batch_add_many.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, C.c_int, vp]
    lib.osqp_batch_add_many.restype = C.c_int
    lib.osqp_batch_poll_many.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
    lib.osqp_batch_poll_many.restype = C.c_int
    lib.osqp_batch_resolve.argtypes = [vp, C.c_int, vp]
    lib.osqp_batch_resolve.restype = C.c_int
    lib.osqp_batch_update_data_mat.argtypes = [vp, C.c_int, vp, vp, C.c_int, vp, vp, C.c_int]
    lib.osqp_batch_update_data_mat.restype = C.c_int
    lib.osqp_batch_step.argtypes = [vp, C.POINTER(C.c_int)]
    lib.osqp_batch_step.restype = C.c_int
    lib.osqp_batch_last_completed.argtypes = [vp, C.POINTER(C.c_int), C.c_int]
    lib.osqp_batch_last_completed.restype = C.c_int
    lib.osqp_batch_poll.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.POINTER(BatchInfo)]
    lib.osqp_batch_poll.restype = C.c_int
    lib.osqp_batch_release.argtypes = [vp, C.c_int]
    lib.osqp_batch_release.restype = C.c_int
    lib.osqp_batch_pending.argtypes = [vp]
    lib.osqp_batch_pending.restype = C.c_int
    lib.osqp_batch_tuning.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.osqp_batch_tuning.restype = C.c_int
    lib.osqp_batch_set_team.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    lib.osqp_batch_set_team.restype = C.c_int
    lib.osqp_batch_counters.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                        C.POINTER(C.c_double)]
    lib.osqp_batch_counters.restype = C.c_int
    lib.osqp_batch_destroy.argtypes = [vp]
    lib.osqp_batch_destroy.restype = C.c_int
    _bound = True
    return lib


def _check_lengths(Ph, Ah, q, l, u, x0, y0):
    """The C entry points read q, x0 over n entries and l, u, y0 over m."""
    n, m = Ph.n, Ah.m
    if Ah.n != n:
        raise ValueError(f"A has {Ah.n} columns, P has {n}")
    if q is None or len(q) != n:
        raise ValueError(f"q must have n = {n} entries")
    for name, v in (("l", l), ("u", u)):
        if v is None or len(v) != m:
            raise ValueError(f"{name} must have m = {m} entries")
    if x0 is not None and len(x0) != n:
        raise ValueError(f"x0 must have n = {n} entries")
    if y0 is not None and len(y0) != m:
        raise ValueError(f"y0 must have m = {m} entries")


class BatchEngine:
    """Many QPs resident on one GPU, advanced together epoch by epoch."""

    def __init__(self, precision=32, max_active=4096, arena_bytes=0, **settings):
        self.lib = _bind(load_library())
        s = T.default_settings()
        T.apply_settings(s, **settings)
        self._settings = s
        self._b = C.c_void_p()
        rc = self.lib.osqp_batch_create(C.byref(self._b), C.addressof(s), int(precision), int(max_active),
                                        int(arena_bytes))
        if rc:
            raise RuntimeError(f"osqp_batch_create failed ({rc})")
        self.precision = precision
        self._dims = {}

    # ---- requests -------------------------------------------------------------
    def add(self, P, q, A, l, u, x0=None, y0=None, resident=False):
        Ph, Ah = T.CscHolder(P), T.CscHolder(A)
        q, l, u = T.as_f64(q), T.as_f64(l), T.as_f64(u)  # l, u are clamped to +-OSQP_INFTY by the engine
        x0, y0 = T.as_f64(x0), T.as_f64(y0)
        _check_lengths(Ph, Ah, q, l, u, x0, y0)
        pid = C.c_int(0)
        rc = self.lib.osqp_batch_add(self._b, Ph.ptr, T.addr(q), Ah.ptr, T.addr(l), T.addr(u), T.addr(x0),
                                     T.addr(y0), 1 if resident else 0, C.byref(pid))
        if rc:
            raise ValueError(f"osqp_batch_add rejected the problem ({rc})")
        self._dims[pid.value] = (Ph.n, Ah.m)
        return pid.value

    def add_many(self, problems, resident=False):
        """Queues a list of problems (dicts with P, q, A, l, u and optional x0,
        y0) in one call; their host-side preparation runs on several threads.
        Returns their ids in order (None for a rejected problem)."""
        k = len(problems)
        if k == 0:
            return []
        packed = _marshal.csc_batch(problems) if _marshal is not None else None
        if packed is not None:
            # every problem already in the engine's form (CSC, int32 / float64,
            # contiguous, exact lengths): its arrays are passed as they are
            buf, base = packed
            ps, ts = C.sizeof(T.OSQPCscMatrix), C.sizeof(C.c_void_p)
            tab = base + 2 * k * ps
            ids = (C.c_int * k)()
            self.lib.osqp_batch_add_many(self._b, k, tab, tab + k * ts, tab + 2 * k * ts, tab + 3 * k * ts,
                                         tab + 4 * k * ts, tab + 5 * k * ts, tab + 6 * k * ts,
                                         1 if resident else 0, ids)
            dims = np.frombuffer(buf, dtype=np.int32, count=2 * k, offset=2 * k * ps + 7 * k * ts)
            idl, nl, ml = list(ids), dims[:k].tolist(), dims[k:].tolist()
            del buf
            self._dims.update((i, (nn, mm)) for i, nn, mm in zip(idl, nl, ml) if i >= 0)
            return [i if i >= 0 else None for i in idl]
        arr = lambda: (C.c_void_p * k)()
        Pp, qp, Ap, lp, up, xp, yp = arr(), arr(), arr(), arr(), arr(), arr(), arr()
        keep = []
        for i, p in enumerate(problems):
            Ph, Ah = T.CscHolder(p["P"]), T.CscHolder(p["A"])
            q, l, u = T.as_f64(p["q"]), T.as_f64(p["l"]), T.as_f64(p["u"])
            x0, y0 = T.as_f64(p.get("x0")), T.as_f64(p.get("y0"))
            _check_lengths(Ph, Ah, q, l, u, x0, y0)
            keep.append((Ph, Ah, q, l, u, x0, y0))
            Pp[i], Ap[i] = C.addressof(Ph.s), C.addressof(Ah.s)
            qp[i], lp[i], up[i], xp[i], yp[i] = T.addr(q), T.addr(l), T.addr(u), T.addr(x0), T.addr(y0)
        ids = (C.c_int * k)()
        self.lib.osqp_batch_add_many(self._b, k, Pp, qp, Ap, lp, up, xp, yp, 1 if resident else 0, ids)
        out = []
        for i, (Ph, Ah, *_rest) in enumerate(keep):
            if ids[i] >= 0:
                self._dims[ids[i]] = (Ph.n, Ah.m)
                out.append(ids[i])
            else:
                out.append(None)
        return out

    def poll_info_many(self, ids):
        """Takes the results of many ended solves in one call, without x and
        y: a dict id -> (status_val, iter, obj_val) for those taken (a
        non-resident problem is released, as by poll)."""
        k = len(ids)
        if k == 0:
            return {}
        idv = np.ascontiguousarray(ids, dtype=np.int32)
        infos = (BatchInfo * k)()
        states = np.empty(k, dtype=np.int32)
        if self.lib.osqp_batch_poll_many(self._b, k, T.addr(idv), None, None, C.addressof(infos),
                                         T.addr(states)) < 0:
            raise ValueError("osqp_batch_poll_many: bad arguments")
        out = {}
        for j in range(k):
            if states[j] == 1:
                inf = infos[j]
                out[int(idv[j])] = (inf.status_val, inf.iter, inf.obj_val)
        return out

    def resolve(self, pid, q=None):
        q = T.as_f64(q)
        if q is not None and pid in self._dims and len(q) != self._dims[pid][0]:
            raise ValueError(f"q must have n = {self._dims[pid][0]} entries")
        rc = self.lib.osqp_batch_resolve(self._b, int(pid), T.addr(q))
        if rc:
            raise RuntimeError(f"osqp_batch_resolve({pid}) failed: not resident, unknown or busy")

    def update_mat(self, pid, Px=None, Px_idx=None, Ax=None, Ax_idx=None):
        """Stages new P (upper-triangle CSC order) / A (CSC order) values of a
        resident, polled problem; the next resolve applies them."""
        Pxn, Axn = T.as_f64(Px), T.as_f64(Ax)
        Pin = None if Px_idx is None else np.ascontiguousarray(Px_idx, dtype=np.int32)
        Ain = None if Ax_idx is None else np.ascontiguousarray(Ax_idx, dtype=np.int32)
        rc = self.lib.osqp_batch_update_data_mat(
            self._b, int(pid), T.addr(Pxn), T.addr(Pin), 0 if Pxn is None else len(Pxn), T.addr(Axn),
            T.addr(Ain), 0 if Axn is None else len(Axn))
        if rc:
            raise ValueError(f"osqp_batch_update_data_mat({pid}) failed ({rc})")

    def step(self):
        """Runs one epoch; returns the ids whose solves ended in it."""
        done = C.c_int(0)
        rc = self.lib.osqp_batch_step(self._b, C.byref(done))
        if rc:
            raise RuntimeError(f"osqp_batch_step failed ({rc})")
        if done.value == 0:
            return []
        ids = (C.c_int * done.value)()
        k = self.lib.osqp_batch_last_completed(self._b, ids, done.value)
        return list(ids[:k])

    def poll(self, pid):
        """Result namespace (x, y, info) once the solve has ended, else None."""
        n, m = self._dims[pid]
        x = np.empty(n)
        y = np.empty(max(m, 1))
        pc = np.empty(max(m, 1))
        dc = np.empty(n)
        info = BatchInfo()
        rc = self.lib.osqp_batch_poll(self._b, int(pid), T.addr(x), T.addr(y), T.addr(pc), T.addr(dc),
                                      C.byref(info))
        if rc < 0:
            raise KeyError(pid)
        if rc == 0:
            return None
        inf = SimpleNamespace(**{f: getattr(info, f) for f, _ in BatchInfo._fields_})
        inf.status = T.STATUS_NAMES.get(inf.status_val, "unknown")
        return SimpleNamespace(x=x, y=y[:m], prim_inf_cert=pc[:m], dual_inf_cert=dc, info=inf)

    def release(self, pid):
        rc = self.lib.osqp_batch_release(self._b, int(pid))
        self._dims.pop(pid, None)
        return rc

    def pending(self):
        return self.lib.osqp_batch_pending(self._b)

    def tuning(self, bucket):
        t, k = C.c_int(0), C.c_int(0)
        if self.lib.osqp_batch_tuning(self._b, int(bucket), C.byref(t), C.byref(k)):
            return None
        return t.value, k.value

    def set_team(self, bucket, team, teams_per_cu):
        if self.lib.osqp_batch_set_team(self._b, int(bucket), int(team), int(teams_per_cu)):
            raise ValueError("unsupported team configuration")

    def counters(self):
        e, i, g = C.c_longlong(0), C.c_longlong(0), C.c_double(0)
        self.lib.osqp_batch_counters(self._b, C.byref(e), C.byref(i), C.byref(g))
        return dict(epochs=e.value, admm_iters=i.value, gpu_s=g.value)

    # ---- convenience ----------------------------------------------------------
    def solve_batch(self, problems, max_epochs=100000):
        """Adds every problem (dicts with P, q, A, l, u), runs epochs until all
        have ended and returns their results in order."""
        ids = self.add_many(problems)
        if any(i is None for i in ids):
            raise ValueError("osqp_batch_add_many rejected a problem")
        out = {}
        for _ in range(max_epochs):
            if not self.pending():
                break
            # a finished solve holds its memory until polled: poll as they end
            for pid in self.step():
                out[pid] = self.poll(pid)
        return [out.get(i) for i in ids]

    def close(self):
        if self._b:
            self.lib.osqp_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
