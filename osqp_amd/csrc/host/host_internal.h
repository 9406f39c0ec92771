/*
 * host_internal.h — solver workspace of the C host core.
 *
 * The host core restates OSQP's solver logic (src/osqp_api.c, src/auxil.c,
 * src/scaling.c) in plain C and reaches the data only through the algebra
 * interface (include/osqp_amd_algebra.h), exactly as the reference core does.
 * The workspace mirrors OSQPWorkspace (include/private/types.h:99-231) minus
 * the code-generation and derivative members this build does not provide.
 */
#ifndef OSQP_AMD_HOST_INTERNAL_H
#define OSQP_AMD_HOST_INTERNAL_H

#include <time.h>

#include "../../../include/osqp_amd.h"
#include "../../../include/osqp_amd_algebra.h"

#define c_max(a, b) (((a) > (b)) ? (a) : (b))
#define c_min(a, b) (((a) < (b)) ? (a) : (b))
#define c_absval(x) (((x) < 0) ? -(x) : (x))

typedef struct {
  OSQPFloat    c;
  OSQPVectorf* D;
  OSQPVectorf* E;
  OSQPFloat    cinv;
  OSQPVectorf* Dinv;
  OSQPVectorf* Einv;
} OSQPScaling;

typedef struct {
  OSQPInt      n;
  OSQPInt      m;
  OSQPMatrix*  P;
  OSQPMatrix*  A;
  OSQPVectorf* q;
  OSQPVectorf* l;
  OSQPVectorf* u;
} OSQPData;

struct OSQPWorkspace_ {
  OSQPData*     data;
  LinSysSolver* linsys_solver;
  OSQPVectorf*  rho_vec;
  OSQPVectorf*  rho_inv_vec;
  OSQPVectori*  constr_type;
  OSQPVectorf *x, *y, *z, *xz_tilde, *xtilde_view, *ztilde_view, *x_prev, *z_prev;
  OSQPVectorf *Ax, *Px, *Aty;
  OSQPFloat    xtPx, qtx, SC, scaled_dual_gap;
  OSQPVectorf *delta_y, *Atdelta_y, *delta_x, *Pdelta_x, *Adelta_x;
  OSQPVectorf *D_temp, *D_temp_A, *E_temp;
  OSQPScaling* scaling;
  OSQPFloat    scaled_prim_res, scaled_dual_res;
  /* the tolerance norms of the last update_info (max of the two / three
   * norms in compute_prim_tol / compute_dual_tol), valid while tol_valid */
  OSQPFloat    tol_prim_norm, tol_dual_norm;
  OSQPInt      tol_valid;
  /* xz_tilde already holds the next iteration's compute_rhs result */
  OSQPInt      rhs_ready;
  OSQPInt      w_ready;   /* ... and the PCG's w for it */
  OSQPFloat    rho_inv;
  struct timespec tic;
  OSQPInt first_run, clear_update_time, rho_update_from_solve;
  OSQPInt rho_updated;
  OSQPFloat last_rel_kkt;
  OSQPInt summary_printed;
  /* polishing (OSQPPolish, include/private/types.h:47-62) */
  OSQPMatrix*  pol_Ared;
  OSQPVectori* pol_active;
  OSQPVectorf *pol_x, *pol_z, *pol_y;
  OSQPInt      pol_n_active;
  OSQPFloat    pol_obj_val, pol_dual_obj_val, pol_duality_gap, pol_prim_res, pol_dual_res;
  OSQPFloat    polish_t;  /* run time of the last polish */
};

/* timer (src/timing_linux.c:29-47) */
void      osqp_amd_tic(OSQPWorkspace* w);
OSQPFloat osqp_amd_toc(OSQPWorkspace* w);

/* auxil (src/auxil.c) */
OSQPFloat compute_rho_estimate(const OSQPSolver* solver);
OSQPInt   adapt_rho(OSQPSolver* solver);
OSQPInt   set_rho_vec(OSQPSolver* solver);
OSQPInt   update_rho_vec(OSQPSolver* solver);
void      swap_vectors(OSQPVectorf** a, OSQPVectorf** b);
void      update_xz_tilde(OSQPSolver* solver, OSQPInt admm_iter);
void      update_x(OSQPSolver* solver);
void      update_z(OSQPSolver* solver);
void      update_y(OSQPSolver* solver);
void      update_xzy(OSQPSolver* solver);
void      compute_obj_val_dual_gap(const OSQPSolver* solver, const OSQPVectorf* x, const OSQPVectorf* y,
                                   OSQPFloat* prim_obj_val, OSQPFloat* dual_obj_val, OSQPFloat* duality_gap);
OSQPInt   has_solution(const OSQPInfo* info);
void      store_solution(OSQPSolver* solver, OSQPSolution* solution);
void      update_info(OSQPSolver* solver, OSQPInt iter);
void      update_info_polish(OSQPSolver* solver);
/* polish (src/polish.c) */
OSQPInt   polish(OSQPSolver* solver);
void      reset_info(OSQPInfo* info);
void      update_status(OSQPInfo* info, OSQPInt status_val);
OSQPInt   check_termination(OSQPSolver* solver, OSQPInt approximate);
OSQPInt   validate_data(const OSQPCscMatrix* P, const OSQPFloat* q, const OSQPCscMatrix* A,
                        const OSQPFloat* l, const OSQPFloat* u, OSQPInt m, OSQPInt n);
OSQPInt   validate_settings(const OSQPSettings* settings, OSQPInt from_setup);

/* scaling (src/scaling.c) */
OSQPInt scale_data(OSQPSolver* solver);
OSQPInt unscale_data(OSQPSolver* solver);
OSQPInt unscale_solution(OSQPVectorf* usolx, OSQPVectorf* usoly, const OSQPVectorf* solx,
                         const OSQPVectorf* soly, OSQPWorkspace* work);

/* errors (src/error.c) */
OSQPInt osqp_amd_error(OSQPInt code, const char* func);
#define osqp_error(code) osqp_amd_error((code), __func__)

#endif
