"""ctypes access to the algebra/hip boundary (include/osqp_amd_algebra.h).

Used by the algebra-boundary tests, which exercise each backend symbol the way
OSQP's lin_alg tests do (tests/lin_alg/*.cpp in the reference).
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _types as T
from .solver import load_library

_ready = False


def lib():
    global _ready
    L = load_library()
    if not _ready:
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        sig = {
            "OSQPVectorf_new": ([vp, i], vp), "OSQPVectori_new": ([vp, i], vp),
            "OSQPVectorf_malloc": ([i], vp), "OSQPVectorf_calloc": ([i], vp), "OSQPVectori_calloc": ([i], vp),
            "OSQPVectorf_copy_new": ([vp], vp), "OSQPVectorf_free": ([vp], None), "OSQPVectori_free": ([vp], None),
            "OSQPVectorf_view": ([vp, i, i], vp), "OSQPVectorf_view_free": ([vp], None),
            "OSQPVectorf_length": ([vp], i), "OSQPVectori_length": ([vp], i),
            "OSQPVectorf_copy": ([vp, vp], None), "OSQPVectorf_from_raw": ([vp, vp], None),
            "OSQPVectorf_to_raw": ([vp, vp], None), "OSQPVectori_to_raw": ([vp, vp], None),
            "OSQPVectori_from_raw": ([vp, vp], None),
            "OSQPVectorf_set_scalar": ([vp, d], None),
            "OSQPVectorf_set_scalar_conditional": ([vp, vp, d, d, d], None),
            "OSQPVectorf_round_to_zero": ([vp, d], None), "OSQPVectorf_mult_scalar": ([vp, d], None),
            "OSQPVectorf_plus": ([vp, vp, vp], None), "OSQPVectorf_minus": ([vp, vp, vp], None),
            "OSQPVectorf_add_scaled": ([vp, d, vp, d, vp], None),
            "OSQPVectorf_add_scaled3": ([vp, d, vp, d, vp, d, vp], None),
            "OSQPVectorf_norm_inf": ([vp], d), "OSQPVectorf_scaled_norm_inf": ([vp, vp], d),
            "OSQPVectorf_norm_inf_diff": ([vp, vp], d), "OSQPVectorf_norm_1": ([vp], d),
            "OSQPVectorf_dot_prod": ([vp, vp], d), "OSQPVectorf_dot_prod_signed": ([vp, vp, i], d),
            "OSQPVectorf_ew_prod": ([vp, vp, vp], None), "OSQPVectorf_all_leq": ([vp, vp], i),
            "OSQPVectorf_ew_bound_vec": ([vp, vp, vp, vp], None),
            "OSQPVectorf_project_polar_reccone": ([vp, vp, vp, d], None),
            "OSQPVectorf_in_reccone": ([vp, vp, vp, d, d], i),
            "OSQPVectorf_ew_reciprocal": ([vp, vp], None), "OSQPVectorf_ew_sqrt": ([vp], None),
            "OSQPVectorf_ew_max_vec": ([vp, vp, vp], None), "OSQPVectorf_ew_min_vec": ([vp, vp, vp], None),
            "OSQPVectorf_ew_bounds_type": ([vp, vp, vp, d, d], i),
            "OSQPVectorf_set_scalar_if_lt": ([vp, vp, d, d], None),
            "OSQPVectorf_set_scalar_if_gt": ([vp, vp, d, d], None),
            "OSQPVectorf_is_eq": ([vp, vp, d], i),
            "OSQPMatrix_new_from_csc": ([vp, i], vp), "OSQPMatrix_free": ([vp], None),
            "OSQPMatrix_get_m": ([vp], i), "OSQPMatrix_get_n": ([vp], i), "OSQPMatrix_get_nz": ([vp], i),
            "OSQPMatrix_update_values": ([vp, vp, vp, i], None), "OSQPMatrix_mult_scalar": ([vp, d], None),
            "OSQPMatrix_lmult_diag": ([vp, vp], None), "OSQPMatrix_rmult_diag": ([vp, vp], None),
            "OSQPMatrix_Axpy": ([vp, vp, vp, d, d], None), "OSQPMatrix_Atxpy": ([vp, vp, vp, d, d], None),
            "OSQPMatrix_col_norm_inf": ([vp, vp], None), "OSQPMatrix_row_norm_inf": ([vp, vp], None),
            "OSQPMatrix_AtDA_extract_diag": ([vp, vp, vp], None), "OSQPMatrix_extract_diag": ([vp, vp], None),
            "OSQPMatrix_submatrix_byrows": ([vp, vp], vp), "OSQPMatrix_is_eq": ([vp, vp, d], i),
            "osqp_algebra_init_libs": ([i], i), "osqp_algebra_free_libs": ([], None),
            "osqp_amd_algebra_sync": ([], None),
            "osqp_amd_red_norm_inf": ([vp, vp], None), "osqp_amd_red_scaled_norm_inf": ([vp, vp, vp], None),
            "osqp_amd_red_dot_prod": ([vp, vp, vp], None), "osqp_amd_red_dot_prod_signed": ([vp, vp, i, vp], None),
            "osqp_amd_red_wait": ([], None),
        }
        for name, (a, r) in sig.items():
            f = getattr(L, name)
            f.argtypes = a
            f.restype = r
        _ready = True
    return L


class Vec:
    """Device FP64 vector handle."""

    def __init__(self, a=None, n=None, handle=None, view=False):
        self.L = lib()
        self.view = view
        if handle is not None:
            self.h, self.n = handle, n
        elif a is not None:
            a = np.ascontiguousarray(a, dtype=np.float64)
            self.n = len(a)
            self.h = self.L.OSQPVectorf_new(a.ctypes.data if self.n else None, self.n)
        else:
            self.n = n
            self.h = self.L.OSQPVectorf_calloc(n)

    def get(self):
        out = np.zeros(self.n)
        if self.n:
            self.L.OSQPVectorf_to_raw(out.ctypes.data, self.h)
        return out

    def free(self):
        if self.h:
            (self.L.OSQPVectorf_view_free if self.view else self.L.OSQPVectorf_free)(self.h)
            self.h = None


class IVec:
    def __init__(self, a=None, n=None):
        self.L = lib()
        if a is not None:
            a = np.ascontiguousarray(a, dtype=np.int32)
            self.n = len(a)
            self.h = self.L.OSQPVectori_new(a.ctypes.data if self.n else None, self.n)
        else:
            self.n = n
            self.h = self.L.OSQPVectori_calloc(n)

    def get(self):
        out = np.zeros(self.n, dtype=np.int32)
        if self.n:
            self.L.OSQPVectori_to_raw(out.ctypes.data, self.h)
        return out

    def free(self):
        if self.h:
            self.L.OSQPVectori_free(self.h)
            self.h = None


class Mat:
    """Device matrix from a scipy matrix (is_triu: P given by its upper triangle)."""

    def __init__(self, M, is_triu=False):
        self.L = lib()
        M = sp.csc_matrix(M)
        if is_triu:
            M = sp.triu(M, format="csc")
        self.holder = T.CscHolder(M)
        self.shape = M.shape
        self.h = self.L.OSQPMatrix_new_from_csc(C.addressof(self.holder.s), 1 if is_triu else 0)

    def free(self):
        if self.h:
            self.L.OSQPMatrix_free(self.h)
            self.h = None


def init(device=0):
    return lib().osqp_algebra_init_libs(device)


def shutdown():
    lib().osqp_algebra_free_libs()
