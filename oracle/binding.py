"""ctypes binding of oracle/liboracle.so (CPU FP64 OSQP restatement).

Test infrastructure only: see oracle/__init__.py.
"""
import ctypes as C
import os

import numpy as np

from osqp_amd._types import (OSQPInfo, OSQPSettings, OSQPSolution, CscHolder, default_settings,
                             apply_settings, dptr, iptr, as_f64, OSQP_DIRECT_SOLVER, OSQP_INFTY)

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def oracle_lib_path():
    return os.path.join(_HERE, "liboracle.so")


def load_oracle():
    global _lib
    if _lib is not None:
        return _lib
    path = oracle_lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.oracle_setup.argtypes = [C.POINTER(vp), vp, vp, vp, vp, vp, C.c_int, C.c_int, vp]
    lib.oracle_setup.restype = C.c_int
    lib.oracle_solve.argtypes = [vp]
    lib.oracle_solve.restype = C.c_int
    lib.oracle_cleanup.argtypes = [vp]
    lib.oracle_info.argtypes = [vp]
    lib.oracle_info.restype = C.POINTER(OSQPInfo)
    lib.oracle_solution.argtypes = [vp]
    lib.oracle_solution.restype = C.POINTER(OSQPSolution)
    lib.oracle_settings.argtypes = [vp]
    lib.oracle_settings.restype = C.POINTER(OSQPSettings)
    lib.oracle_update_data_vec.argtypes = [vp, vp, vp, vp]
    lib.oracle_update_data_vec.restype = C.c_int
    lib.oracle_warm_start.argtypes = [vp, vp, vp]
    lib.oracle_warm_start.restype = C.c_int
    lib.oracle_update_rho.argtypes = [vp, C.c_double]
    lib.oracle_update_rho.restype = C.c_int
    lib.oracle_update_settings.argtypes = [vp, vp]
    lib.oracle_update_settings.restype = C.c_int
    lib.oracle_update_data_mat.argtypes = [vp, vp, vp, C.c_int, vp, vp, C.c_int]
    lib.oracle_update_data_mat.restype = C.c_int
    lib.oracle_total_pcg_iters.argtypes = [vp]
    lib.oracle_total_pcg_iters.restype = C.c_long
    lib.oracle_get_scaling.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.oracle_get_scaled_matrices.argtypes = [vp, vp, vp]
    for name in ("oracle_csc_Axpy", "oracle_csc_Atxpy"):
        f = getattr(lib, name)
        f.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_double, C.c_double]
    lib.oracle_csc_Axpy_sym_triu.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_double, C.c_double]
    for name in ("oracle_csc_col_norm_inf", "oracle_csc_row_norm_inf"):
        getattr(lib, name).argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
    lib.oracle_csc_row_norm_inf_sym_triu.argtypes = [C.c_int, vp, vp, vp, vp]
    lib.oracle_csc_AtDA_extract_diag.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp]
    lib.oracle_csc_to_csr.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp]
    lib.oracle_triu_to_full_csr.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp]
    lib.oracle_kkt_solve.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, C.c_double, vp, vp]
    lib.oracle_kkt_solve.restype = C.c_int
    lib.oracle_set_default_settings.argtypes = [vp]
    _lib = lib
    return lib


def clip_bounds(v):
    if v is None:
        return None
    return np.clip(as_f64(v), -OSQP_INFTY, OSQP_INFTY)


class OracleSolver:
    """Mirror of the osqp_* API over the CPU oracle (same argument meaning)."""

    def __init__(self):
        self.lib = load_oracle()
        self.h = C.c_void_p()

    def setup(self, P, q, A, l, u, **settings):
        import scipy.sparse as sp
        lib = self.lib
        P = sp.triu(sp.csc_matrix(P), format="csc")
        A = sp.csc_matrix(A)
        self.n = P.shape[0]
        self.m = A.shape[0]
        self._P = CscHolder(P)
        self._A = CscHolder(A)
        self._q = as_f64(q)
        self._l = clip_bounds(l) if self.m else np.zeros(1)
        self._u = clip_bounds(u) if self.m else np.zeros(1)
        st = default_settings()
        st.linsys_solver = OSQP_DIRECT_SOLVER
        apply_settings(st, **settings)
        self._settings = st
        rc = lib.oracle_setup(C.byref(self.h), C.addressof(self._P.s), self._q.ctypes.data,
                              C.addressof(self._A.s), self._l.ctypes.data, self._u.ctypes.data,
                              self.m, self.n, C.addressof(st))
        return rc

    def solve(self):
        return self.lib.oracle_solve(self.h)

    @property
    def info(self):
        return self.lib.oracle_info(self.h).contents

    @property
    def settings(self):
        return self.lib.oracle_settings(self.h).contents

    @property
    def x(self):
        s = self.lib.oracle_solution(self.h).contents
        return np.ctypeslib.as_array(s.x, shape=(self.n,)).copy()

    @property
    def y(self):
        s = self.lib.oracle_solution(self.h).contents
        return np.ctypeslib.as_array(s.y, shape=(self.m,)).copy() if self.m else np.zeros(0)

    @property
    def prim_inf_cert(self):
        s = self.lib.oracle_solution(self.h).contents
        return np.ctypeslib.as_array(s.prim_inf_cert, shape=(self.m,)).copy() if self.m else np.zeros(0)

    @property
    def dual_inf_cert(self):
        s = self.lib.oracle_solution(self.h).contents
        return np.ctypeslib.as_array(s.dual_inf_cert, shape=(self.n,)).copy()

    def update_data_vec(self, q=None, l=None, u=None):
        self._qn, self._ln, self._un = as_f64(q), clip_bounds(l), clip_bounds(u)
        return self.lib.oracle_update_data_vec(
            self.h, self._qn.ctypes.data if q is not None else None,
            self._ln.ctypes.data if l is not None else None,
            self._un.ctypes.data if u is not None else None)

    def update_data_mat(self, Px=None, Px_idx=None, Ax=None, Ax_idx=None, Px_n0=False, Ax_n0=False):
        """osqp_update_data_mat; Px_n0 / Ax_n0 pass a count of 0 (all values)."""
        self._keep = [as_f64(Px), None if Px_idx is None else np.ascontiguousarray(Px_idx, np.int32),
                      as_f64(Ax), None if Ax_idx is None else np.ascontiguousarray(Ax_idx, np.int32)]
        Pxn, Pin, Axn, Ain = self._keep
        return self.lib.oracle_update_data_mat(
            self.h, Pxn.ctypes.data if Pxn is not None else None,
            Pin.ctypes.data if Pin is not None else None, 0 if Pxn is None or Px_n0 else len(Pxn),
            Axn.ctypes.data if Axn is not None else None,
            Ain.ctypes.data if Ain is not None else None, 0 if Axn is None or Ax_n0 else len(Axn))

    def warm_start(self, x=None, y=None):
        self._wx, self._wy = as_f64(x), as_f64(y)
        return self.lib.oracle_warm_start(self.h, self._wx.ctypes.data if x is not None else None,
                                          self._wy.ctypes.data if y is not None else None)

    def update_rho(self, rho):
        return self.lib.oracle_update_rho(self.h, float(rho))

    def update_settings(self, **kw):
        st = OSQPSettings()
        C.pointer(st)[0] = self.settings
        apply_settings(st, **kw)
        return self.lib.oracle_update_settings(self.h, C.addressof(st))

    def total_pcg_iters(self):
        return self.lib.oracle_total_pcg_iters(self.h)

    def scaling(self):
        n, m = self.n, self.m
        D, E, c = np.zeros(n), np.zeros(max(m, 1)), C.c_double()
        q, l, u = np.zeros(n), np.zeros(max(m, 1)), np.zeros(max(m, 1))
        self.lib.oracle_get_scaling(self.h, D.ctypes.data, E.ctypes.data, C.addressof(c),
                                    q.ctypes.data, l.ctypes.data, u.ctypes.data)
        return D, E[:m], c.value, q, l[:m], u[:m]

    def scaled_matrices(self):
        Px = np.zeros(max(len(self._P.x), 1))
        Ax = np.zeros(max(len(self._A.x), 1))
        self.lib.oracle_get_scaled_matrices(self.h, Px.ctypes.data, Ax.ctypes.data)
        return Px, Ax

    def __del__(self):
        try:
            if self.h:
                self.lib.oracle_cleanup(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass
