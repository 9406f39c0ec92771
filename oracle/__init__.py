"""CPU oracle for the MI355X OSQP engine — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  The product library never loads it.
"""
from .binding import OracleSolver, load_oracle, oracle_lib_path  # noqa: F401
