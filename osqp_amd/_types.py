"""ctypes mirror of include/osqp_amd.h (the reference's public OSQP types).

Layouts follow include/public/osqp_api_types.h:37-196 with OSQPFloat=double and
OSQPInt=int32.  This module loads no library; it is shared by the product
binding (osqp_amd.solver) and by the test tooling.
"""
import ctypes as C

import numpy as np

OSQPInt = C.c_int
OSQPFloat = C.c_double

# include/public/osqp_api_constants.h:28-40
OSQP_SOLVED = 1
OSQP_SOLVED_INACCURATE = 2
OSQP_PRIMAL_INFEASIBLE = 3
OSQP_PRIMAL_INFEASIBLE_INACCURATE = 4
OSQP_DUAL_INFEASIBLE = 5
OSQP_DUAL_INFEASIBLE_INACCURATE = 6
OSQP_MAX_ITER_REACHED = 7
OSQP_TIME_LIMIT_REACHED = 8
OSQP_NON_CVX = 9
OSQP_SIGINT = 10
OSQP_UNSOLVED = 11

STATUS_NAMES = {
    1: "solved", 2: "solved inaccurate", 3: "primal infeasible",
    4: "primal infeasible inaccurate", 5: "dual infeasible",
    6: "dual infeasible inaccurate", 7: "maximum iterations reached",
    8: "run time limit reached", 9: "problem non convex", 10: "interrupted",
    11: "unsolved",
}

OSQP_DIRECT_SOLVER = 1
OSQP_INDIRECT_SOLVER = 2

# include/public/osqp_api_constants.h:74-88
OSQP_NO_ERROR = 0
OSQP_DATA_VALIDATION_ERROR = 1
OSQP_SETTINGS_VALIDATION_ERROR = 2
OSQP_LINSYS_SOLVER_INIT_ERROR = 3
OSQP_NONCVX_ERROR = 4
OSQP_MEM_ALLOC_ERROR = 5
OSQP_WORKSPACE_NOT_INIT_ERROR = 6
OSQP_ALGEBRA_LOAD_ERROR = 7

OSQP_INFTY = 1e30
# include/public/osqp_api_constants.h:193 — the float value of the integer
# 0x7fc00000, not an IEEE NaN.
OSQP_NAN = float(0x7FC00000)


class OSQPCscMatrix(C.Structure):
    _fields_ = [
        ("m", OSQPInt), ("n", OSQPInt),
        # OSQPInt* p, OSQPInt* i, OSQPFloat* x (held as addresses)
        ("p", C.c_void_p), ("i", C.c_void_p), ("x", C.c_void_p),
        ("nzmax", OSQPInt), ("nz", OSQPInt), ("owned", OSQPInt),
    ]


class OSQPSettings(C.Structure):
    _fields_ = [
        ("device", OSQPInt), ("linsys_solver", C.c_int),
        ("allocate_solution", OSQPInt), ("verbose", OSQPInt), ("profiler_level", OSQPInt),
        ("warm_starting", OSQPInt), ("scaling", OSQPInt), ("polishing", OSQPInt),
        ("rho", OSQPFloat), ("rho_is_vec", OSQPInt), ("sigma", OSQPFloat), ("alpha", OSQPFloat),
        ("cg_max_iter", OSQPInt), ("cg_tol_reduction", OSQPInt), ("cg_tol_fraction", OSQPFloat),
        ("cg_precond", C.c_int),
        ("adaptive_rho", OSQPInt), ("adaptive_rho_interval", OSQPInt),
        ("adaptive_rho_fraction", OSQPFloat), ("adaptive_rho_tolerance", OSQPFloat),
        ("max_iter", OSQPInt), ("eps_abs", OSQPFloat), ("eps_rel", OSQPFloat),
        ("eps_prim_inf", OSQPFloat), ("eps_dual_inf", OSQPFloat),
        ("scaled_termination", OSQPInt), ("check_termination", OSQPInt),
        ("check_dualgap", OSQPInt), ("time_limit", OSQPFloat),
        ("delta", OSQPFloat), ("polish_refine_iter", OSQPInt),
    ]


class OSQPInfo(C.Structure):
    _fields_ = [
        ("status", C.c_char * 32), ("status_val", OSQPInt), ("status_polish", OSQPInt),
        ("obj_val", OSQPFloat), ("dual_obj_val", OSQPFloat), ("prim_res", OSQPFloat),
        ("dual_res", OSQPFloat), ("duality_gap", OSQPFloat), ("iter", OSQPInt),
        ("rho_updates", OSQPInt), ("rho_estimate", OSQPFloat), ("setup_time", OSQPFloat),
        ("solve_time", OSQPFloat), ("update_time", OSQPFloat), ("polish_time", OSQPFloat),
        ("run_time", OSQPFloat), ("primdual_int", OSQPFloat), ("rel_kkt_error", OSQPFloat),
    ]


class OSQPSolution(C.Structure):
    _fields_ = [
        ("x", C.POINTER(OSQPFloat)), ("y", C.POINTER(OSQPFloat)),
        ("prim_inf_cert", C.POINTER(OSQPFloat)), ("dual_inf_cert", C.POINTER(OSQPFloat)),
    ]


class OSQPSolver(C.Structure):
    _fields_ = [
        ("settings", C.POINTER(OSQPSettings)), ("solution", C.POINTER(OSQPSolution)),
        ("info", C.POINTER(OSQPInfo)), ("work", C.c_void_p),
    ]


def default_settings():
    """Values of osqp_set_default_settings (src/osqp_api.c:283-327), non-CUDA build."""
    s = OSQPSettings()
    s.device = 0
    s.linsys_solver = OSQP_INDIRECT_SOLVER
    s.allocate_solution = 1
    s.verbose = 0
    s.profiler_level = 0
    s.warm_starting = 1
    s.scaling = 10
    s.polishing = 0
    s.rho = 0.1
    s.rho_is_vec = 1
    s.sigma = 1e-6
    s.alpha = 1.6
    s.cg_max_iter = 20
    s.cg_tol_reduction = 10
    s.cg_tol_fraction = 0.15
    s.cg_precond = 1
    s.adaptive_rho = 1
    s.adaptive_rho_interval = 50
    s.adaptive_rho_fraction = 0.4
    s.adaptive_rho_tolerance = 5.0
    s.max_iter = 4000
    s.eps_abs = 1e-3
    s.eps_rel = 1e-3
    s.eps_prim_inf = 1e-4
    s.eps_dual_inf = 1e-4
    s.scaled_termination = 0
    s.check_termination = 25
    s.check_dualgap = 1
    s.time_limit = 1e10
    s.delta = 1e-6
    s.polish_refine_iter = 3
    return s


def apply_settings(s, **kw):
    for k, v in kw.items():
        if not hasattr(s, k):
            raise AttributeError(f"unknown setting {k}")
        setattr(s, k, v)
    return s


class CscHolder:
    """Keeps numpy buffers alive behind an OSQPCscMatrix."""

    def __init__(self, M):
        import scipy.sparse as sp
        # a CSC matrix with sorted int32 indices is used as it is (no copy)
        if not sp.issparse(M) or M.format != "csc":
            M = sp.csc_matrix(M)
        if not M.has_sorted_indices:
            M = M.sorted_indices()
        self.m, self.n = M.shape
        self.p = np.ascontiguousarray(M.indptr, dtype=np.int32)
        self.i = np.ascontiguousarray(M.indices, dtype=np.int32)
        self.x = np.ascontiguousarray(M.data, dtype=np.float64)
        if self.x.size == 0:
            self.i = np.zeros(1, np.int32)
            self.x = np.zeros(1, np.float64)
        self.s = OSQPCscMatrix(self.m, self.n, addr(self.p), addr(self.i), addr(self.x), int(M.nnz), -1, 0)

    @property
    def ptr(self):
        return C.byref(self.s)


def addr(a):
    """Address of a numpy array's data as an int (None for None): the cheap
    form for arguments declared c_void_p."""
    return None if a is None else a.__array_interface__["data"][0]


def dptr(a):
    return a.ctypes.data_as(C.POINTER(OSQPFloat)) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(C.POINTER(OSQPInt)) if a is not None else None


def as_f64(a):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.float64)
