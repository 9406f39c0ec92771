"""Python front end of libosqp_amd.so (ctypes over the C ABI in include/osqp_amd.h).

The class mirrors the method set of OSQP's Python interface (setup / solve /
update / warm_start / update_settings / update_rho / cold_start), so code and
tests written against OSQP read the same.  All numerical work happens in the
HIP library; this module only marshals arrays.  It raises if the library is
missing — there is no CPU fallback.
"""
import ctypes as C
import os
import threading
from types import SimpleNamespace

import numpy as np
import scipy.sparse as sp

from . import _types as T

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libosqp_amd.so")
_lib = None
_lock = threading.Lock()


def load_library():
    """Loads libosqp_amd.so (built in-tree by osqp_amd.build)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libosqp_amd.so not built: {LIB_PATH} (run `python -m osqp_amd.build`)")
        lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        lib.osqp_setup.argtypes = [C.POINTER(C.POINTER(T.OSQPSolver)), vp, vp, vp, vp, vp, C.c_int, C.c_int, vp]
        lib.osqp_setup.restype = C.c_int
        lib.osqp_solve.argtypes = [C.POINTER(T.OSQPSolver)]
        lib.osqp_solve.restype = C.c_int
        lib.osqp_cleanup.argtypes = [C.POINTER(T.OSQPSolver)]
        lib.osqp_cleanup.restype = C.c_int
        lib.osqp_update_data_vec.argtypes = [C.POINTER(T.OSQPSolver), vp, vp, vp]
        lib.osqp_update_data_vec.restype = C.c_int
        lib.osqp_update_data_mat.argtypes = [C.POINTER(T.OSQPSolver), vp, vp, C.c_int, vp, vp, C.c_int]
        lib.osqp_update_data_mat.restype = C.c_int
        lib.osqp_warm_start.argtypes = [C.POINTER(T.OSQPSolver), vp, vp]
        lib.osqp_warm_start.restype = C.c_int
        lib.osqp_cold_start.argtypes = [C.POINTER(T.OSQPSolver)]
        lib.osqp_update_rho.argtypes = [C.POINTER(T.OSQPSolver), C.c_double]
        lib.osqp_update_rho.restype = C.c_int
        lib.osqp_update_settings.argtypes = [C.POINTER(T.OSQPSolver), vp]
        lib.osqp_update_settings.restype = C.c_int
        lib.osqp_set_default_settings.argtypes = [vp]
        lib.osqp_get_solution.argtypes = [C.POINTER(T.OSQPSolver), C.POINTER(T.OSQPSolution)]
        lib.osqp_get_solution.restype = C.c_int
        lib.osqp_capabilities.restype = C.c_int
        lib.osqp_version.restype = C.c_char_p
        lib.osqp_error_message.restype = C.c_char_p
        lib.osqp_error_message.argtypes = [C.c_int]
        lib.osqp_amd_set_pcg_precision.argtypes = [C.c_int]
        lib.osqp_amd_set_pcg_precision.restype = C.c_int
        lib.osqp_amd_get_pcg_precision.restype = C.c_int
        lib.osqp_amd_solver_stats.argtypes = [C.POINTER(T.OSQPSolver), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                              C.POINTER(C.c_double)]
        lib.osqp_amd_solver_stats.restype = C.c_int
        lib.osqp_amd_apply_operator.argtypes = [C.POINTER(T.OSQPSolver), C.c_int, vp, vp]
        lib.osqp_amd_apply_operator.restype = C.c_int
        lib.osqp_amd_time_pcg_kernels.argtypes = [C.POINTER(T.OSQPSolver), C.c_int] + [C.POINTER(C.c_double)] * 4
        lib.osqp_amd_time_pcg_kernels.restype = C.c_int
        lib.osqp_amd_time_fp64_kernels.argtypes = [C.POINTER(T.OSQPSolver), C.c_int] + [C.POINTER(C.c_double)] * 4
        lib.osqp_amd_time_fp64_kernels.restype = C.c_int
        lib.osqp_amd_pcg_bytes.argtypes = [C.POINTER(T.OSQPSolver), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.osqp_amd_pcg_bytes.restype = C.c_int
        lib.osqp_amd_check_bytes.argtypes = [C.POINTER(T.OSQPSolver), C.POINTER(C.c_double)]
        lib.osqp_amd_check_bytes.restype = C.c_int
        lib.osqp_algebra_device_name.argtypes = [C.c_char_p, C.c_int]
        lib.osqp_amd_last_error.restype = C.c_int
        lib.osqp_amd_profiler_counts.argtypes = [C.POINTER(C.c_longlong)] * 3
        lib.osqp_amd_profiler_counts.restype = C.c_int
        lib.osqp_amd_matrix_export.argtypes = [vp] + [vp] * 11
        lib.osqp_amd_matrix_export.restype = C.c_int
        lib.osqp_amd_solver_matrices.argtypes = [C.POINTER(T.OSQPSolver), C.POINTER(vp), C.POINTER(vp)]
        lib.osqp_amd_solver_matrices.restype = C.c_int
        lib.osqp_amd_solver_scaling.argtypes = [C.POINTER(T.OSQPSolver), vp, vp, C.POINTER(C.c_double), vp, vp, vp]
        lib.osqp_amd_solver_scaling.restype = C.c_int
        _lib = lib
        return lib


def profiler_counts():
    """roctx ranges pushed / popped and events marked so far in this process."""
    lib = load_library()
    a, b, c = C.c_longlong(), C.c_longlong(), C.c_longlong()
    lib.osqp_amd_profiler_counts(C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def matrix_export(handle, m, n, is_triu):
    """Host copies of an OSQPMatrix's backend structures (see
    osqp_amd_matrix_export in include/osqp_amd_algebra.h)."""
    lib = load_library()
    nc, ni = C.c_int(), C.c_int()
    lib.osqp_amd_matrix_export(handle, C.addressof(nc), C.addressof(ni), *([None] * 9))
    rows = n if is_triu else m
    nnz, nin = nc.value, ni.value
    out = dict(rp=np.zeros(rows + 1, np.int32), ci=np.zeros(max(nnz, 1), np.int32), val=np.zeros(max(nnz, 1)),
               csc2csr=np.zeros(max(nin, 1), np.int32))
    if is_triu:
        out.update(mirror=np.zeros(max(nin, 1), np.int32), diag=np.zeros(max(n, 1), np.int32))
    else:
        out.update(trp=np.zeros(n + 1, np.int32), tci=np.zeros(max(nin, 1), np.int32), tval=np.zeros(max(nin, 1)))
    g = lambda k: out[k].ctypes.data if k in out else None
    lib.osqp_amd_matrix_export(handle, None, None, g("rp"), g("ci"), g("val"), g("trp"), g("tci"), g("tval"),
                               g("csc2csr"), g("mirror"), g("diag"))
    for k, ln in (("ci", nnz), ("val", nnz), ("csc2csr", nin), ("mirror", nin), ("diag", n), ("tci", nin),
                  ("tval", nin)):
        if k in out:
            out[k] = out[k][:ln]
    return out


def default_settings():
    lib = load_library()
    s = T.OSQPSettings()
    lib.osqp_set_default_settings(C.addressof(s))
    s.verbose = 0
    return s


def device_name():
    lib = load_library()
    buf = C.create_string_buffer(256)
    lib.osqp_algebra_device_name(buf, 256)
    return buf.value.decode()


class OSQP:
    """One QP instance on the GPU (osqp_setup ... osqp_cleanup).

    precision: 64 -> plain FP64 backend; 32 -> FP32-PCG with FP64 residuals.
    """

    def __init__(self, precision=64):
        self.lib = load_library()
        self.precision = precision
        self._solver = C.POINTER(T.OSQPSolver)()
        self._keep = []

    # -- setup ----------------------------------------------------------------
    def setup(self, P, q, A, l, u, **settings):
        lib = self.lib
        if P is None:
            raise ValueError("P is required")
        P = sp.triu(sp.csc_matrix(P), format="csc")
        A = sp.csc_matrix(A) if A is not None else sp.csc_matrix((0, P.shape[0]))
        self.n, self.m = P.shape[0], A.shape[0]
        self._P = T.CscHolder(P)
        self._A = T.CscHolder(A)
        self._q = T.as_f64(q)
        self._l = np.clip(T.as_f64(l), -T.OSQP_INFTY, T.OSQP_INFTY) if self.m else np.zeros(1)
        self._u = np.clip(T.as_f64(u), -T.OSQP_INFTY, T.OSQP_INFTY) if self.m else np.zeros(1)
        st = default_settings()
        T.apply_settings(st, **settings)
        self._settings = st
        if lib.osqp_amd_set_pcg_precision(int(self.precision)):
            raise ValueError("precision must be 32 or 64")
        rc = lib.osqp_setup(C.byref(self._solver), C.addressof(self._P.s), self._q.ctypes.data,
                            C.addressof(self._A.s), self._l.ctypes.data, self._u.ctypes.data,
                            self.m, self.n, C.addressof(st))
        return rc

    # -- solve ----------------------------------------------------------------
    def solve(self):
        rc = self.lib.osqp_solve(self._solver)
        sol = self._solver.contents.solution.contents
        n, m = self.n, self.m
        x = np.ctypeslib.as_array(sol.x, shape=(n,)).copy()
        y = np.ctypeslib.as_array(sol.y, shape=(m,)).copy() if m else np.zeros(0)
        pic = np.ctypeslib.as_array(sol.prim_inf_cert, shape=(m,)).copy() if m else np.zeros(0)
        dic = np.ctypeslib.as_array(sol.dual_inf_cert, shape=(n,)).copy()
        return SimpleNamespace(x=x, y=y, prim_inf_cert=pic, dual_inf_cert=dic, info=self.info, rc=rc)

    @property
    def info(self):
        i = self._solver.contents.info.contents
        return SimpleNamespace(**{k: (getattr(i, k).decode() if k == "status" else getattr(i, k))
                                  for k, _ in T.OSQPInfo._fields_})

    @property
    def settings(self):
        return self._solver.contents.settings.contents

    # -- updates --------------------------------------------------------------
    def update(self, q=None, l=None, u=None, Px=None, Px_idx=None, Ax=None, Ax_idx=None):
        rc = 0
        if q is not None or l is not None or u is not None:
            qq = T.as_f64(q)
            ll = None if l is None else np.clip(T.as_f64(l), -T.OSQP_INFTY, T.OSQP_INFTY)
            uu = None if u is None else np.clip(T.as_f64(u), -T.OSQP_INFTY, T.OSQP_INFTY)
            self._keep = [qq, ll, uu]
            rc = self.lib.osqp_update_data_vec(self._solver, None if qq is None else qq.ctypes.data,
                                               None if ll is None else ll.ctypes.data,
                                               None if uu is None else uu.ctypes.data)
            if rc:
                return rc
        if Px is not None or Ax is not None:
            Pxn, Axn = T.as_f64(Px), T.as_f64(Ax)
            Pin = None if Px_idx is None else np.ascontiguousarray(Px_idx, dtype=np.int32)
            Ain = None if Ax_idx is None else np.ascontiguousarray(Ax_idx, dtype=np.int32)
            self._keep += [Pxn, Axn, Pin, Ain]
            rc = self.lib.osqp_update_data_mat(
                self._solver, None if Pxn is None else Pxn.ctypes.data, None if Pin is None else Pin.ctypes.data,
                0 if Pxn is None else len(Pxn), None if Axn is None else Axn.ctypes.data,
                None if Ain is None else Ain.ctypes.data, 0 if Axn is None else len(Axn))
        return rc

    def warm_start(self, x=None, y=None):
        xx, yy = T.as_f64(x), T.as_f64(y)
        self._keep = [xx, yy]
        return self.lib.osqp_warm_start(self._solver, None if xx is None else xx.ctypes.data,
                                        None if yy is None else yy.ctypes.data)

    def cold_start(self):
        self.lib.osqp_cold_start(self._solver)

    def update_rho(self, rho):
        return self.lib.osqp_update_rho(self._solver, float(rho))

    def update_settings(self, **kw):
        st = T.OSQPSettings()
        C.pointer(st)[0] = self.settings
        T.apply_settings(st, **kw)
        return self.lib.osqp_update_settings(self._solver, C.addressof(st))

    def stats(self):
        it, sv, ms = C.c_int(), C.c_int(), C.c_double()
        self.lib.osqp_amd_solver_stats(self._solver, C.byref(it), C.byref(sv), C.byref(ms))
        return SimpleNamespace(pcg_iters=it.value, linsys_solves=sv.value, solve_ms=ms.value)

    def apply_operator(self, which, x):
        """The linear system's operator on x through its own kernels:
        which 0: K x in the CG precision; 1: K x with the FP64-vector operator
        of the defect correction; 2: A x (z = A x of the defect correction).
        K = P + sigma I + A' diag(rho) A on the solver's scaled data."""
        n, m = self.n, self.m
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(m if which == 2 else n)
        if self.lib.osqp_amd_apply_operator(self._solver, int(which), x.ctypes.data, y.ctypes.data):
            raise RuntimeError("osqp_amd_apply_operator failed")
        return y

    def time_kernels(self, reps=20):
        """Mean ms and algorithmic bytes of one launch of each CG operator kernel."""
        a, k, ba, bk = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        rc = self.lib.osqp_amd_time_pcg_kernels(self._solver, reps, C.byref(a), C.byref(k), C.byref(ba), C.byref(bk))
        if rc:
            raise RuntimeError("osqp_amd_time_pcg_kernels failed")
        return dict(ms_av=a.value, ms_kv=k.value, bytes_av=ba.value, bytes_kv=bk.value)

    def time_fp64_kernels(self, reps=20):
        """Mean ms and algorithmic bytes of one launch of the defect
        correction's FP64 panel SpMVs (A x and [P A'] [x; w]); None when the
        operators are not panelled."""
        a, k, ba, bk = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        if self.lib.osqp_amd_time_fp64_kernels(self._solver, reps, C.byref(a), C.byref(k), C.byref(ba), C.byref(bk)):
            return None
        return dict(ms_a=a.value, ms_k=k.value, bytes_a=ba.value, bytes_k=bk.value)

    def pcg_bytes(self):
        """Algorithmic HBM bytes of one CG iteration and of one linear solve's
        work outside the CG loop."""
        a, b = C.c_double(), C.c_double()
        if self.lib.osqp_amd_pcg_bytes(self._solver, C.byref(a), C.byref(b)):
            raise RuntimeError("osqp_amd_pcg_bytes failed")
        return dict(per_cg_iter=a.value, per_solve=b.value)

    def check_bytes(self):
        """Algorithmic HBM bytes of one termination check (the products A x,
        P x, A' y and the vector passes of update_info)."""
        b = C.c_double()
        if self.lib.osqp_amd_check_bytes(self._solver, C.byref(b)):
            raise RuntimeError("osqp_amd_check_bytes failed")
        return b.value

    def scaling(self):
        """Ruiz scaling after setup: D, E, c and the scaled q, l, u (None
        when the solver runs without scaling)."""
        n, m = self.n, self.m
        D, E, q, l, u = np.zeros(n), np.zeros(max(m, 1)), np.zeros(n), np.zeros(max(m, 1)), np.zeros(max(m, 1))
        c = C.c_double()
        if self.lib.osqp_amd_solver_scaling(self._solver, D.ctypes.data, E.ctypes.data, C.byref(c), q.ctypes.data,
                                            l.ctypes.data, u.ctypes.data):
            return None
        return D, E[:m], c.value, q, l[:m], u[:m]

    def structures(self):
        """The backend's CSR structures of P (full symmetric) and A, with
        their maps from the input CSC entries (osqp_amd_matrix_export)."""
        Ph, Ah = C.c_void_p(), C.c_void_p()
        self.lib.osqp_amd_solver_matrices(self._solver, C.byref(Ph), C.byref(Ah))
        return dict(P=matrix_export(Ph.value, self.n, self.n, True), A=matrix_export(Ah.value, self.m, self.n, False))

    def close(self):
        if self._solver:
            self.lib.osqp_cleanup(self._solver)
            self._solver = C.POINTER(T.OSQPSolver)()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
