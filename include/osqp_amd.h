/*
 * osqp_amd.h — public C ABI of the MI355X OSQP engine.
 *
 * The types, constants and entry points below keep the layout and meaning of
 * OSQP's public API so that a C host written against OSQP compiles and links
 * against libosqp_amd.so unchanged:
 *   types      include/public/osqp_api_types.h:37-196  (OSQPCscMatrix, OSQPSettings,
 *              OSQPInfo, OSQPSolution, OSQPSolver)
 *   constants  include/public/osqp_api_constants.h:28-222
 *   functions  include/public/osqp_api_functions.h:40-456
 *
 * Build contract: OSQPFloat = double, OSQPInt = int (32-bit).  Every entry point
 * is plain C with plain pointers; no device types cross this boundary.  Input
 * arrays may be host or device pointers where OSQP's CUDA backend accepts both
 * (vector from_raw / to_raw), matrices are always read from host memory
 * (include/public/osqp_api_functions.h:267, src/auxil.c:992-1014).
 */
#ifndef OSQP_AMD_H
#define OSQP_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int    OSQPInt;
typedef double OSQPFloat;

/* ---- constants (include/public/osqp_api_constants.h) -------------------- */

enum osqp_capabilities_type {
  OSQP_CAPABILITY_DIRECT_SOLVER   = 0x01,
  OSQP_CAPABILITY_INDIRECT_SOLVER = 0x02,
  OSQP_CAPABILITY_CODEGEN         = 0x04,
  OSQP_CAPABILITY_UPDATE_MATRICES = 0x08,
  OSQP_CAPABILITY_DERIVATIVES     = 0x10
};

enum osqp_status_type {
  OSQP_SOLVED = 1,
  OSQP_SOLVED_INACCURATE,
  OSQP_PRIMAL_INFEASIBLE,
  OSQP_PRIMAL_INFEASIBLE_INACCURATE,
  OSQP_DUAL_INFEASIBLE,
  OSQP_DUAL_INFEASIBLE_INACCURATE,
  OSQP_MAX_ITER_REACHED,
  OSQP_TIME_LIMIT_REACHED,
  OSQP_NON_CVX,
  OSQP_SIGINT,
  OSQP_UNSOLVED
};

enum osqp_polish_status_type {
  OSQP_POLISH_LINSYS_ERROR = -2,
  OSQP_POLISH_FAILED = -1,
  OSQP_POLISH_NOT_PERFORMED = 0,
  OSQP_POLISH_SUCCESS = 1,
  OSQP_POLISH_NO_ACTIVE_SET_FOUND = 2
};

enum osqp_linsys_solver_type {
  OSQP_UNKNOWN_SOLVER = 0,
  OSQP_DIRECT_SOLVER,
  OSQP_INDIRECT_SOLVER
};

typedef enum {
  OSQP_NO_PRECONDITIONER = 0,
  OSQP_DIAGONAL_PRECONDITIONER
} osqp_precond_type;

enum osqp_error_type {
  OSQP_NO_ERROR = 0,
  OSQP_DATA_VALIDATION_ERROR = 1,
  OSQP_SETTINGS_VALIDATION_ERROR,
  OSQP_LINSYS_SOLVER_INIT_ERROR,
  OSQP_NONCVX_ERROR,
  OSQP_MEM_ALLOC_ERROR,
  OSQP_WORKSPACE_NOT_INIT_ERROR,
  OSQP_ALGEBRA_LOAD_ERROR,
  OSQP_FOPEN_ERROR,
  OSQP_CODEGEN_DEFINES_ERROR,
  OSQP_DATA_NOT_INITIALIZED,
  OSQP_FUNC_NOT_IMPLEMENTED,
  OSQP_LAST_ERROR_PLACE
};

#define OSQP_VERBOSE        (1)
#define OSQP_WARM_STARTING  (1)
#define OSQP_SCALING        (10)
#define OSQP_POLISHING      (0)
#define OSQP_RHO            (0.1)
#define OSQP_SIGMA          (1E-06)
#define OSQP_ALPHA          (1.6)
#define OSQP_RHO_MIN        (1e-06)
#define OSQP_RHO_MAX        (1e06)
#define OSQP_RHO_TOL        (1e-04)
#define OSQP_RHO_EQ_OVER_RHO_INEQ (1e03)
#define OSQP_RHO_IS_VEC     (1)
#define OSQP_CG_MAX_ITER      (20)
#define OSQP_CG_TOL_REDUCTION (10)
#define OSQP_CG_TOL_FRACTION  (0.15)
#define OSQP_ADAPTIVE_RHO_UPDATE_DISABLED   (0)
#define OSQP_ADAPTIVE_RHO_UPDATE_ITERATIONS (1)
#define OSQP_ADAPTIVE_RHO_UPDATE_TIME       (2)
#define OSQP_ADAPTIVE_RHO_UPDATE_KKT_ERROR  (3)
#define _OSQP_ADAPTIVE_RHO_UPDATE_LAST_VALUE (4)
#define OSQP_ADAPTIVE_RHO_UPDATE_DEFAULT (OSQP_ADAPTIVE_RHO_UPDATE_ITERATIONS)
#define OSQP_ADAPTIVE_RHO_INTERVAL  (50)
#define OSQP_ADAPTIVE_RHO_TOLERANCE (5.0)
#define OSQP_ADAPTIVE_RHO_FRACTION  (0.4)
#define OSQP_ADAPTIVE_RHO_MULTIPLE_TERMINATION (4)
#define OSQP_ADAPTIVE_RHO_FIXED (100)
#define OSQP_MAX_ITER           (4000)
#define OSQP_EPS_ABS            (1E-3)
#define OSQP_EPS_REL            (1E-3)
#define OSQP_EPS_PRIM_INF       (1E-4)
#define OSQP_EPS_DUAL_INF       (1E-4)
#define OSQP_SCALED_TERMINATION (0)
#define OSQP_TIME_LIMIT         (1e10)
#define OSQP_CHECK_DUALGAP      (1)
#define OSQP_CHECK_TERMINATION  (25)
#define OSQP_DELTA              (1E-6)
#define OSQP_POLISH_REFINE_ITER (3)
#define OSQP_NULL 0
/* Not an IEEE NaN: the reference defines it as the float value of the bit
 * pattern 0x7fc00000 read as an integer (osqp_api_constants.h:193). */
#define OSQP_NAN   ((OSQPFloat)0x7fc00000UL)
#define OSQP_INFTY ((OSQPFloat)1e30)
#define OSQP_DIVISION_TOL (1.0 / OSQP_INFTY)
#define OSQP_PRINT_INTERVAL 200
#define OSQP_MIN_SCALING   (1e-04)
#define OSQP_MAX_SCALING   (1e+04)
#define OSQP_CG_TOL_MIN    (1E-7)
#define OSQP_CG_POLISH_TOL (1e-5)
#define OSQP_ZERO_DEADZONE (1e-15)

/* ---- types (include/public/osqp_api_types.h) --------------------------- */

typedef struct {
  OSQPInt    m;
  OSQPInt    n;
  OSQPInt   *p;
  OSQPInt   *i;
  OSQPFloat *x;
  OSQPInt    nzmax;
  OSQPInt    nz;
  OSQPInt    owned;
} OSQPCscMatrix;

typedef struct {
  OSQPInt device;
  enum osqp_linsys_solver_type linsys_solver;
  OSQPInt allocate_solution;
  OSQPInt verbose;
  OSQPInt profiler_level;
  OSQPInt warm_starting;
  OSQPInt scaling;
  OSQPInt polishing;
  OSQPFloat rho;
  OSQPInt   rho_is_vec;
  OSQPFloat sigma;
  OSQPFloat alpha;
  OSQPInt   cg_max_iter;
  OSQPInt   cg_tol_reduction;
  OSQPFloat cg_tol_fraction;
  osqp_precond_type cg_precond;
  OSQPInt   adaptive_rho;
  OSQPInt   adaptive_rho_interval;
  OSQPFloat adaptive_rho_fraction;
  OSQPFloat adaptive_rho_tolerance;
  OSQPInt   max_iter;
  OSQPFloat eps_abs;
  OSQPFloat eps_rel;
  OSQPFloat eps_prim_inf;
  OSQPFloat eps_dual_inf;
  OSQPInt   scaled_termination;
  OSQPInt   check_termination;
  OSQPInt   check_dualgap;
  OSQPFloat time_limit;
  OSQPFloat delta;
  OSQPInt   polish_refine_iter;
} OSQPSettings;

typedef struct {
  char      status[32];
  OSQPInt   status_val;
  OSQPInt   status_polish;
  OSQPFloat obj_val;
  OSQPFloat dual_obj_val;
  OSQPFloat prim_res;
  OSQPFloat dual_res;
  OSQPFloat duality_gap;
  OSQPInt   iter;
  OSQPInt   rho_updates;
  OSQPFloat rho_estimate;
  OSQPFloat setup_time;
  OSQPFloat solve_time;
  OSQPFloat update_time;
  OSQPFloat polish_time;
  OSQPFloat run_time;
  OSQPFloat primdual_int;
  OSQPFloat rel_kkt_error;
} OSQPInfo;

typedef struct {
  OSQPFloat* x;
  OSQPFloat* y;
  OSQPFloat* prim_inf_cert;
  OSQPFloat* dual_inf_cert;
} OSQPSolution;

typedef struct OSQPWorkspace_ OSQPWorkspace;

typedef struct {
  OSQPSettings*  settings;
  OSQPSolution*  solution;
  OSQPInfo*      info;
  OSQPWorkspace* work;
} OSQPSolver;

extern const char* OSQP_STATUS_MESSAGE[];
extern const char* OSQP_ERROR_MESSAGE[];

/* ---- entry points (include/public/osqp_api_functions.h) ---------------- */

/* osqp_api_functions.h:211 */
OSQPInt     osqp_capabilities(void);
/* osqp_api_functions.h:218 */
const char* osqp_version(void);
/* osqp_api_functions.h:227 */
const char* osqp_error_message(OSQPInt error_flag);
/* osqp_api_functions.h:237 */
void        osqp_get_dimensions(OSQPSolver* solver, OSQPInt* m, OSQPInt* n);
/* osqp_api_functions.h:414 ; defaults src/osqp_api.c:283-327 */
void        osqp_set_default_settings(OSQPSettings* settings);
/* osqp_api_functions.h:155,164 */
OSQPSettings* OSQPSettings_new(void);
void          OSQPSettings_free(OSQPSettings* settings);
/* osqp_api_functions.h:40,58 */
OSQPCscMatrix* OSQPCscMatrix_new(OSQPInt m, OSQPInt n, OSQPInt nzmax,
                                 OSQPFloat* x, OSQPInt* i, OSQPInt* p);
void           OSQPCscMatrix_free(OSQPCscMatrix* mat);
/* osqp_api_functions.h:132 */
void OSQPCscMatrix_set_data(OSQPCscMatrix* M, OSQPInt m, OSQPInt n, OSQPInt nzmax,
                            OSQPFloat* x, OSQPInt* i, OSQPInt* p);

/* osqp_api_functions.h:267 ; src/osqp_api.c:332-631 */
OSQPInt osqp_setup(OSQPSolver** solverp,
                   const OSQPCscMatrix* P, const OSQPFloat* q,
                   const OSQPCscMatrix* A, const OSQPFloat* l, const OSQPFloat* u,
                   OSQPInt m, OSQPInt n, const OSQPSettings* settings);
/* osqp_api_functions.h:295 ; src/osqp_api.c:636-1038 */
OSQPInt osqp_solve(OSQPSolver* solver);
/* osqp_api_functions.h:306 ; src/osqp_api.c:1041 */
OSQPInt osqp_get_solution(OSQPSolver* solver, OSQPSolution* solution);
/* osqp_api_functions.h:317 ; src/osqp_api.c:1057-1174 */
OSQPInt osqp_cleanup(OSQPSolver* solver);
/* osqp_api_functions.h:344 ; src/osqp_api.c:1260-1293 */
OSQPInt osqp_warm_start(OSQPSolver* solver, const OSQPFloat* x, const OSQPFloat* y);
/* osqp_api_functions.h:352 ; src/osqp_api.c:1296-1304 */
void    osqp_cold_start(OSQPSolver* solver);
/* osqp_api_functions.h:362 ; src/osqp_api.c:1184-1257 */
OSQPInt osqp_update_data_vec(OSQPSolver* solver, const OSQPFloat* q_new,
                             const OSQPFloat* l_new, const OSQPFloat* u_new);
/* osqp_api_functions.h:388 ; src/osqp_api.c:1309-1409 */
OSQPInt osqp_update_data_mat(OSQPSolver* solver,
                             const OSQPFloat* Px_new, const OSQPInt* Px_new_idx, OSQPInt P_new_n,
                             const OSQPFloat* Ax_new, const OSQPInt* Ax_new_idx, OSQPInt A_new_n);
/* osqp_api_functions.h:437 ; src/osqp_api.c:1474-1530 */
OSQPInt osqp_update_settings(OSQPSolver* solver, const OSQPSettings* new_settings);
/* osqp_api_functions.h:454 ; src/osqp_api.c:1412-1464 */
OSQPInt osqp_update_rho(OSQPSolver* solver, OSQPFloat rho_new);

/* ---- MI355X extensions (no counterpart in the reference) ---------------
 * PCG arithmetic: 64 = FP64 values and vectors everywhere (the plain
 * backend); 32 = FP32 values/vectors inside the PCG with compensated dot
 * products, while residuals, termination, certificates and the rho ratio stay
 * FP64.  Applies to solvers created after the call; default 64.           */
OSQPInt osqp_amd_set_pcg_precision(OSQPInt bits);
OSQPInt osqp_amd_get_pcg_precision(void);
/* Counters of the last osqp_solve on this solver: total PCG iterations and
 * the GPU time (ms) spent between the first and last kernel of the solve. */
OSQPInt osqp_amd_solver_stats(OSQPSolver* solver, OSQPInt* pcg_iters_total,
                              OSQPInt* linsys_solves, double* gpu_ms);

/* Times the CG operator kernels of this solver's linear system (see
 * osqp_amd_linsys_time_kernels in osqp_amd_algebra.h). */
OSQPInt osqp_amd_time_pcg_kernels(OSQPSolver* solver, OSQPInt reps, double* ms_av, double* ms_kv,
                                  double* bytes_av, double* bytes_kv);

/* Applies this solver's linear-system operator to the host vector x (see
 * osqp_amd_linsys_apply in osqp_amd_algebra.h; the scaled data). */
OSQPInt osqp_amd_apply_operator(OSQPSolver* solver, OSQPInt which, const OSQPFloat* x, OSQPFloat* y);

/* Times the FP64 panel SpMVs of the PCG's defect correction (see
 * osqp_amd_linsys_time_fp64_kernels in osqp_amd_algebra.h). */
OSQPInt osqp_amd_time_fp64_kernels(OSQPSolver* solver, OSQPInt reps, double* ms_a, double* ms_k,
                                   double* bytes_a, double* bytes_k);

/* Algorithmic HBM bytes of this solver's PCG work: one CG iteration, and the
 * per-solve work around the CG (residual, correction, z = A x) on the cached
 * path (see osqp_amd_linsys_bytes in osqp_amd_algebra.h). */
OSQPInt osqp_amd_pcg_bytes(OSQPSolver* solver, double* per_cg_iter, double* per_solve);

/* Algorithmic HBM bytes of one termination check (update_info: the products
 * A x, P x, A' y and the vector passes of the residuals, objective terms and
 * tolerance norms).  No counterpart in the reference: a measurement helper. */
OSQPInt osqp_amd_check_bytes(OSQPSolver* solver, double* bytes);

/* Number of roctx ranges pushed and popped and of events marked by the
 * profiler annotations (settings->profiler_level 1 or 2; the sections of
 * src/profilers.c:4-28).  For tests of the annotation backend. */
OSQPInt osqp_amd_profiler_counts(long long* pushes, long long* pops, long long* marks);

#ifdef __cplusplus
}
#endif

#endif /* OSQP_AMD_H */
