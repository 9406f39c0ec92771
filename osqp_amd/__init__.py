"""osqp_amd — OSQP's ADMM QP engine on AMD MI355X (gfx950).

Native code: libosqp_amd.so (C host core + hand-written HIP kernels), built
in-tree by `python -m osqp_amd.build`.  This package is a thin ctypes front end.
"""
from ._types import (OSQP_SOLVED, OSQP_SOLVED_INACCURATE, OSQP_PRIMAL_INFEASIBLE,  # noqa: F401
                     OSQP_PRIMAL_INFEASIBLE_INACCURATE, OSQP_DUAL_INFEASIBLE,
                     OSQP_DUAL_INFEASIBLE_INACCURATE, OSQP_MAX_ITER_REACHED, OSQP_TIME_LIMIT_REACHED, OSQP_SIGINT,
                     OSQP_NON_CVX, OSQP_UNSOLVED, OSQP_INFTY, OSQP_NAN, STATUS_NAMES)
from .solver import OSQP, load_library, default_settings, device_name, LIB_PATH  # noqa: F401
