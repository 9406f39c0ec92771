/*
 * auxil.c — ADMM steps, residuals, termination, infeasibility detection, rho
 * adaptation and Ruiz scaling of the host core, expressed as calls into the
 * algebra/hip backend.
 *
 * Each function restates its reference counterpart (file:line given) with the
 * same sequence of algebra calls, so the iterates on the GPU follow the
 * reference's arithmetic step by step.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "host_internal.h"

const char* OSQP_STATUS_MESSAGE[] = {"",
                                     "solved",
                                     "solved inaccurate",
                                     "primal infeasible",
                                     "primal infeasible inaccurate",
                                     "dual infeasible",
                                     "dual infeasible inaccurate",
                                     "maximum iterations reached",
                                     "run time limit reached",
                                     "problem non convex",
                                     "interrupted",
                                     "unsolved"};

const char* OSQP_ERROR_MESSAGE[] = {"",
                                    "Problem data validation.",
                                    "Solver settings validation.",
                                    "Linear system solver initialization.",
                                    "KKT matrix LDL factorization.",
                                    "Memory allocation.",
                                    "Solver workspace not initialized.",
                                    "Error loading algebra library.",
                                    "Error opening file for writing.",
                                    "Error setting codegen defines.",
                                    "Data not initialized.",
                                    "Function not implemented."};

/* src/error.c:25-33 */
OSQPInt osqp_amd_error(OSQPInt code, const char* func) {
  if (code > 0 && code < OSQP_LAST_ERROR_PLACE)
    fprintf(stderr, "ERROR in %s: %s\n", func, OSQP_ERROR_MESSAGE[code]);
  return code;
}

void osqp_amd_tic(OSQPWorkspace* w) { clock_gettime(CLOCK_MONOTONIC, &w->tic); }

OSQPFloat osqp_amd_toc(OSQPWorkspace* w) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (OSQPFloat)(t.tv_sec - w->tic.tv_sec) + 1e-9 * (OSQPFloat)(t.tv_nsec - w->tic.tv_nsec);
}

/* auxil.c:14-47 */
OSQPFloat compute_rho_estimate(const OSQPSolver* solver) {
  OSQPWorkspace* work = solver->work;
  OSQPFloat prim_res = work->scaled_prim_res, dual_res = work->scaled_dual_res;
  OSQPFloat prim_res_norm, dual_res_norm, temp, rho_estimate, nz, nAx, nq, nAty, nPx;
  /* the five norms queued, one host round trip */
  osqp_amd_red_norm_inf(work->z, &nz);
  osqp_amd_red_norm_inf(work->Ax, &nAx);
  osqp_amd_red_norm_inf(work->data->q, &nq);
  osqp_amd_red_norm_inf(work->Aty, &nAty);
  osqp_amd_red_norm_inf(work->Px, &nPx);
  osqp_amd_red_wait();
  prim_res_norm = nz;
  temp = nAx;
  prim_res_norm = c_max(prim_res_norm, temp);
  prim_res /= (prim_res_norm + OSQP_DIVISION_TOL);
  dual_res_norm = nq;
  temp = nAty;
  dual_res_norm = c_max(dual_res_norm, temp);
  temp = nPx;
  dual_res_norm = c_max(dual_res_norm, temp);
  dual_res /= (dual_res_norm + OSQP_DIVISION_TOL);
  rho_estimate = solver->settings->rho * sqrt(prim_res / dual_res);
  return c_min(c_max(rho_estimate, OSQP_RHO_MIN), OSQP_RHO_MAX);
}

/* auxil.c:49-74 */
OSQPInt adapt_rho(OSQPSolver* solver) {
  OSQPInt exitflag = 0;
  OSQPFloat rho_new = compute_rho_estimate(solver);
  solver->work->rhs_ready = 0;  /* compute_rhs uses rho_inv */
  solver->info->rho_estimate = rho_new;
  if ((rho_new > solver->settings->rho * solver->settings->adaptive_rho_tolerance) ||
      (rho_new < solver->settings->rho / solver->settings->adaptive_rho_tolerance)) {
    exitflag = osqp_update_rho(solver, rho_new);
    solver->info->rho_updates += 1;
    solver->work->rho_updated = 1;
  }
  return exitflag;
}

/* auxil.c:76-105 */
OSQPInt set_rho_vec(OSQPSolver* solver) {
  OSQPSettings* settings = solver->settings;
  OSQPWorkspace* work = solver->work;
  OSQPInt changed;
  settings->rho = c_min(c_max(settings->rho, OSQP_RHO_MIN), OSQP_RHO_MAX);
  changed = OSQPVectorf_ew_bounds_type(work->constr_type, work->data->l, work->data->u, OSQP_RHO_TOL,
                                       OSQP_INFTY * OSQP_MIN_SCALING);
  OSQPVectorf_set_scalar_conditional(work->rho_vec, work->constr_type, OSQP_RHO_MIN, settings->rho,
                                     OSQP_RHO_EQ_OVER_RHO_INEQ * settings->rho);
  OSQPVectorf_ew_reciprocal(work->rho_inv_vec, work->rho_vec);
  return changed;
}

/* auxil.c:107-122 */
OSQPInt update_rho_vec(OSQPSolver* solver) {
  OSQPWorkspace* work = solver->work;
  OSQPInt exitflag = 0;
  if (set_rho_vec(solver) == 1)
    exitflag = work->linsys_solver->update_rho_vec(work->linsys_solver, work->rho_vec, solver->settings->rho);
  return exitflag;
}

void swap_vectors(OSQPVectorf** a, OSQPVectorf** b) {
  OSQPVectorf* t = *b;
  *b = *a;
  *a = t;
}

/* auxil.c:136-158 */
static void compute_rhs(OSQPSolver* solver) {
  OSQPWorkspace* work = solver->work;
  OSQPSettings* settings = solver->settings;
  if (work->rhs_ready) {
    /* written by the previous iteration's fused update (same arithmetic),
     * with the PCG's w for it when w_ready */
    work->rhs_ready = 0;
    osqp_amd_linsys_set_w_given(work->linsys_solver, work->w_ready);
    work->w_ready = 0;
    return;
  }
  osqp_amd_linsys_set_w_given(work->linsys_solver, 0);
  if (settings->rho_is_vec) {
    /* add_scaled(xt, sigma, x_prev, -1, q); ew_prod(zt, rho_inv, y);
     * add_scaled(zt, -1, zt, 1, z_prev) — one fused pass */
    osqp_amd_admm_rhs(work->xtilde_view, work->ztilde_view, settings->sigma, work->x_prev, work->data->q,
                      work->rho_inv_vec, work->y, work->z_prev);
  } else {
    OSQPVectorf_add_scaled(work->xtilde_view, settings->sigma, work->x_prev, -1., work->data->q);
    OSQPVectorf_add_scaled(work->ztilde_view, 1.0, work->z_prev, -work->rho_inv, work->y);
  }
}

/* auxil.c:160-170 */
void update_xz_tilde(OSQPSolver* solver, OSQPInt admm_iter) {
  compute_rhs(solver);
  solver->work->linsys_solver->solve(solver->work->linsys_solver, solver->work->xz_tilde, admm_iter);
}

/* auxil.c:172-229 for the rho-vector case: update_x, update_z, update_y in
 * one fused pass (same element arithmetic), which also computes the next
 * iteration's compute_rhs (auxil.c:136-158) from the new x, z, y: the next
 * compute_rhs is then skipped unless rho, sigma or q change first (adapt_rho
 * and every solve's start clear rhs_ready). */
void update_xzy(OSQPSolver* solver) {
  OSQPWorkspace* work = solver->work;
  OSQPFloat* w = osqp_amd_linsys_w_target(work->linsys_solver);
  osqp_amd_admm_update(solver->settings->alpha, work->x, work->delta_x, work->xtilde_view, work->x_prev, work->z,
                       work->y, work->delta_y, work->ztilde_view, work->z_prev, work->rho_vec, work->rho_inv_vec,
                       work->data->l, work->data->u, 1, solver->settings->sigma, work->data->q, w);
  work->rhs_ready = 1;
  work->w_ready = w != NULL;
}

/* auxil.c:172-184 */
void update_x(OSQPSolver* solver) {
  OSQPWorkspace* work = solver->work;
  OSQPFloat a = solver->settings->alpha;
  OSQPVectorf_add_scaled(work->x, a, work->xtilde_view, (1.0 - a), work->x_prev);
  OSQPVectorf_minus(work->delta_x, work->x, work->x_prev);
}

/* auxil.c:186-208 */
void update_z(OSQPSolver* solver) {
  OSQPWorkspace* work = solver->work;
  OSQPFloat a = solver->settings->alpha;
  if (solver->settings->rho_is_vec) {
    OSQPVectorf_ew_prod(work->z, work->rho_inv_vec, work->y);
    OSQPVectorf_add_scaled3(work->z, 1., work->z, a, work->ztilde_view, (1.0 - a), work->z_prev);
  } else {
    OSQPVectorf_add_scaled3(work->z, a, work->ztilde_view, (1.0 - a), work->z_prev, work->rho_inv, work->y);
  }
  OSQPVectorf_ew_bound_vec(work->z, work->z, work->data->l, work->data->u);
}

/* auxil.c:210-229 */
void update_y(OSQPSolver* solver) {
  OSQPWorkspace* work = solver->work;
  OSQPFloat a = solver->settings->alpha;
  OSQPVectorf_add_scaled3(work->delta_y, a, work->ztilde_view, (1.0 - a), work->z_prev, -1.0, work->z);
  if (solver->settings->rho_is_vec) OSQPVectorf_ew_prod(work->delta_y, work->delta_y, work->rho_vec);
  else OSQPVectorf_mult_scalar(work->delta_y, solver->settings->rho);
  OSQPVectorf_plus(work->y, work->y, work->delta_y);
}

/* auxil.c:231-285, in two parts: the dots are queued (obj_terms_queue), then
 * combined once they have arrived (obj_terms_finish). */
typedef struct {
  OSQPFloat quad, lin, sup_u, sup_l;
} ObjTerms;

/* scratch: an m-vector for the projection of y (the reference uses z_prev) */
static void obj_terms_queue(const OSQPSolver* solver, const OSQPVectorf* x, const OSQPVectorf* y,
                            OSQPVectorf* scratch, ObjTerms* t) {
  OSQPWorkspace* work = solver->work;
  osqp_amd_red_dot_prod(work->Px, x, &t->quad);
  osqp_amd_red_dot_prod(work->data->q, x, &t->lin);
  OSQPVectorf_copy(scratch, y);
  OSQPVectorf_project_polar_reccone(scratch, work->data->l, work->data->u, OSQP_INFTY * OSQP_MIN_SCALING);
  OSQPVectorf_round_to_zero(scratch, OSQP_ZERO_DEADZONE);
  osqp_amd_red_dot_prod_signed(work->data->u, scratch, +1, &t->sup_u);
  osqp_amd_red_dot_prod_signed(work->data->l, scratch, -1, &t->sup_l);
}

static void obj_terms_finish(const OSQPSolver* solver, const ObjTerms* t, OSQPFloat* prim_obj_val,
                             OSQPFloat* dual_obj_val, OSQPFloat* duality_gap) {
  OSQPWorkspace* work = solver->work;
  const OSQPFloat quad = t->quad, lin = t->lin;
  OSQPFloat sup = t->sup_u;
  sup += t->sup_l;
  *prim_obj_val = 0.5 * quad + lin;
  *dual_obj_val = -0.5 * quad - sup;
  work->scaled_dual_gap = quad + lin + sup;
  if (solver->settings->scaling) {
    *prim_obj_val *= work->scaling->cinv;
    *dual_obj_val *= work->scaling->cinv;
    *duality_gap = work->scaling->cinv * work->scaled_dual_gap;
  } else {
    *duality_gap = work->scaled_dual_gap;
  }
  work->xtPx = quad;
  work->qtx = lin;
  work->SC = sup;
}

void compute_obj_val_dual_gap(const OSQPSolver* solver, const OSQPVectorf* x, const OSQPVectorf* y,
                              OSQPFloat* prim_obj_val, OSQPFloat* dual_obj_val, OSQPFloat* duality_gap) {
  ObjTerms t;
  obj_terms_queue(solver, x, y, solver->work->z_prev, &t);
  osqp_amd_red_wait();
  obj_terms_finish(solver, &t, prim_obj_val, dual_obj_val, duality_gap);
}

static OSQPInt unscaled_termination(const OSQPSolver* s) {
  return s->settings->scaling && !s->settings->scaled_termination;
}

/* auxil.c:287-306 */
static OSQPFloat compute_duality_gap_tol(const OSQPSolver* solver, OSQPFloat eps_abs, OSQPFloat eps_rel) {
  OSQPWorkspace* work = solver->work;
  OSQPFloat r = c_absval(work->xtPx);
  r = c_max(r, c_absval(work->qtx));
  r = c_max(r, c_absval(work->SC));
  if (unscaled_termination(solver)) r = work->scaling->cinv * r;
  return eps_abs + eps_rel * r;
}

/* auxil.c:308-332 */
static OSQPFloat compute_prim_res(OSQPSolver* solver, const OSQPVectorf* x, const OSQPVectorf* z) {
  OSQPWorkspace* work = solver->work;
  OSQPMatrix_Axpy(work->data->A, x, work->Ax, 1.0, 0.0);
  OSQPVectorf_minus(work->z_prev, work->Ax, z);
  work->scaled_prim_res = OSQPVectorf_norm_inf(work->z_prev);
  if (unscaled_termination(solver)) return OSQPVectorf_scaled_norm_inf(work->scaling->Einv, work->z_prev);
  return work->scaled_prim_res;
}

/* auxil.c:334-369 */
static OSQPFloat compute_prim_tol(const OSQPSolver* solver, OSQPFloat eps_abs, OSQPFloat eps_rel) {
  OSQPWorkspace* work = solver->work;
  OSQPFloat a, b;
  if (work->tol_valid) return eps_abs + eps_rel * work->tol_prim_norm;
  if (unscaled_termination(solver)) {
    a = OSQPVectorf_scaled_norm_inf(work->scaling->Einv, work->z);
    b = OSQPVectorf_scaled_norm_inf(work->scaling->Einv, work->Ax);
  } else {
    a = OSQPVectorf_norm_inf(work->z);
    b = OSQPVectorf_norm_inf(work->Ax);
  }
  return eps_abs + eps_rel * c_max(a, b);
}

/* auxil.c:371-410 */
static OSQPFloat compute_dual_res(OSQPSolver* solver, const OSQPVectorf* x, const OSQPVectorf* y) {
  OSQPWorkspace* work = solver->work;
  OSQPVectorf_copy(work->x_prev, work->data->q);
  OSQPMatrix_Axpy(work->data->P, x, work->Px, 1.0, 0.0);
  OSQPVectorf_plus(work->x_prev, work->x_prev, work->Px);
  if (work->data->m) {
    OSQPMatrix_Atxpy(work->data->A, y, work->Aty, 1.0, 0.0);
    OSQPVectorf_plus(work->x_prev, work->x_prev, work->Aty);
  }
  work->scaled_dual_res = OSQPVectorf_norm_inf(work->x_prev);
  if (unscaled_termination(solver))
    return work->scaling->cinv * OSQPVectorf_scaled_norm_inf(work->scaling->Dinv, work->x_prev);
  return work->scaled_dual_res;
}

/* auxil.c:412-458 */
static OSQPFloat compute_dual_tol(const OSQPSolver* solver, OSQPFloat eps_abs, OSQPFloat eps_rel) {
  OSQPWorkspace* work = solver->work;
  OSQPFloat r, t;
  if (work->tol_valid) return eps_abs + eps_rel * work->tol_dual_norm;
  if (unscaled_termination(solver)) {
    r = OSQPVectorf_scaled_norm_inf(work->scaling->Dinv, work->data->q);
    t = OSQPVectorf_scaled_norm_inf(work->scaling->Dinv, work->Aty);
    r = c_max(r, t);
    t = OSQPVectorf_scaled_norm_inf(work->scaling->Dinv, work->Px);
    r = c_max(r, t);
    r *= work->scaling->cinv;
  } else {
    r = OSQPVectorf_norm_inf(work->data->q);
    t = OSQPVectorf_norm_inf(work->Aty);
    r = c_max(r, t);
    t = OSQPVectorf_norm_inf(work->Px);
    r = c_max(r, t);
  }
  return eps_abs + eps_rel * r;
}

/* auxil.c:460-512 */
static OSQPInt is_primal_infeasible(OSQPSolver* solver, OSQPFloat eps_prim_inf) {
  OSQPWorkspace* work = solver->work;
  OSQPFloat norm_delta_y, ineq_lhs;
  OSQPVectorf_project_polar_reccone(work->delta_y, work->data->l, work->data->u, OSQP_INFTY * OSQP_MIN_SCALING);
  if (unscaled_termination(solver)) {
    OSQPVectorf_ew_prod(work->Adelta_x, work->scaling->E, work->delta_y);
    norm_delta_y = OSQPVectorf_norm_inf(work->Adelta_x);
  } else {
    norm_delta_y = OSQPVectorf_norm_inf(work->delta_y);
  }
  if (norm_delta_y > OSQP_DIVISION_TOL) {
    ineq_lhs = OSQPVectorf_dot_prod_signed(work->data->u, work->delta_y, +1);
    ineq_lhs += OSQPVectorf_dot_prod_signed(work->data->l, work->delta_y, -1);
    if (ineq_lhs < 0.0) {
      OSQPMatrix_Atxpy(work->data->A, work->delta_y, work->Atdelta_y, 1.0, 0.0);
      if (unscaled_termination(solver)) OSQPVectorf_ew_prod(work->Atdelta_y, work->Atdelta_y, work->scaling->Dinv);
      return OSQPVectorf_norm_inf(work->Atdelta_y) < eps_prim_inf * norm_delta_y;
    }
  }
  return 0;
}

/* auxil.c:514-588 */
static OSQPInt is_dual_infeasible(OSQPSolver* solver, OSQPFloat eps_dual_inf) {
  OSQPWorkspace* work = solver->work;
  OSQPFloat norm_delta_x, cost_scaling;
  if (unscaled_termination(solver)) {
    norm_delta_x = OSQPVectorf_scaled_norm_inf(work->scaling->D, work->delta_x);
    cost_scaling = work->scaling->c;
  } else {
    norm_delta_x = OSQPVectorf_norm_inf(work->delta_x);
    cost_scaling = 1.0;
  }
  if (norm_delta_x > OSQP_DIVISION_TOL) {
    if (OSQPVectorf_dot_prod(work->data->q, work->delta_x) < 0.0) {
      OSQPMatrix_Axpy(work->data->P, work->delta_x, work->Pdelta_x, 1.0, 0.0);
      if (unscaled_termination(solver)) OSQPVectorf_ew_prod(work->Pdelta_x, work->Pdelta_x, work->scaling->Dinv);
      if (OSQPVectorf_norm_inf(work->Pdelta_x) < cost_scaling * eps_dual_inf * norm_delta_x) {
        OSQPMatrix_Axpy(work->data->A, work->delta_x, work->Adelta_x, 1.0, 0.0);
        if (unscaled_termination(solver)) OSQPVectorf_ew_prod(work->Adelta_x, work->Adelta_x, work->scaling->Einv);
        return OSQPVectorf_in_reccone(work->Adelta_x, work->data->l, work->data->u, OSQP_INFTY * OSQP_MIN_SCALING,
                                      eps_dual_inf * norm_delta_x);
      }
    }
  }
  return 0;
}

/* auxil.c:590-596 */
OSQPInt has_solution(const OSQPInfo* info) {
  return (info->status_val != OSQP_PRIMAL_INFEASIBLE) && (info->status_val != OSQP_PRIMAL_INFEASIBLE_INACCURATE) &&
         (info->status_val != OSQP_DUAL_INFEASIBLE) && (info->status_val != OSQP_DUAL_INFEASIBLE_INACCURATE) &&
         (info->status_val != OSQP_NON_CVX);
}

/* auxil.c:598-674 */
void store_solution(OSQPSolver* solver, OSQPSolution* solution) {
  OSQPInfo* info = solver->info;
  OSQPWorkspace* work = solver->work;
  OSQPFloat norm_vec;
  if (!solution) return;
  if (has_solution(info)) {
    if (solver->settings->scaling) {
      unscale_solution(work->x_prev, work->z_prev, work->x, work->y, work);
      OSQPVectorf_to_raw(solution->x, work->x_prev);
      OSQPVectorf_to_raw(solution->y, work->z_prev);
    } else {
      OSQPVectorf_to_raw(solution->x, work->x);
      OSQPVectorf_to_raw(solution->y, work->y);
    }
    OSQPVectorf_set_scalar(work->delta_y, OSQP_NAN);
    OSQPVectorf_set_scalar(work->delta_x, OSQP_NAN);
    OSQPVectorf_to_raw(solution->prim_inf_cert, work->delta_y);
    OSQPVectorf_to_raw(solution->dual_inf_cert, work->delta_x);
  } else {
    OSQPVectorf_set_scalar(work->x, OSQP_NAN);
    OSQPVectorf_set_scalar(work->y, OSQP_NAN);
    OSQPVectorf_to_raw(solution->x, work->x);
    OSQPVectorf_to_raw(solution->y, work->y);
    osqp_cold_start(solver);
    if ((info->status_val == OSQP_PRIMAL_INFEASIBLE) || (info->status_val == OSQP_PRIMAL_INFEASIBLE_INACCURATE)) {
      norm_vec = OSQPVectorf_norm_inf(work->delta_y);
      OSQPVectorf_mult_scalar(work->delta_y, 1. / norm_vec);
      OSQPVectorf_to_raw(solution->prim_inf_cert, work->delta_y);
      OSQPVectorf_set_scalar(work->delta_x, OSQP_NAN);
      OSQPVectorf_to_raw(solution->dual_inf_cert, work->delta_x);
    }
    if ((info->status_val == OSQP_DUAL_INFEASIBLE) || (info->status_val == OSQP_DUAL_INFEASIBLE_INACCURATE)) {
      norm_vec = OSQPVectorf_norm_inf(work->delta_x);
      OSQPVectorf_mult_scalar(work->delta_x, 1. / norm_vec);
      OSQPVectorf_to_raw(solution->dual_inf_cert, work->delta_x);
      OSQPVectorf_set_scalar(work->delta_y, OSQP_NAN);
      OSQPVectorf_to_raw(solution->prim_inf_cert, work->delta_y);
    }
  }
}

/* auxil.c:676-762 (non-polishing branch) */
/* The same vector operations as compute_prim_res, compute_dual_res,
 * compute_obj_val_dual_gap, compute_prim_tol and compute_dual_tol, in the
 * same order, with their thirteen reductions queued and run together (one
 * launch and one host round trip instead of thirteen; the same arithmetic,
 * so the same bits).  Queued reductions read their vectors when they run,
 * so the projection of y for the objective terms goes to Adelta_x (free
 * here; the infeasibility checks overwrite it before reading) instead of
 * z_prev, which still holds A x - z.  The tolerance norms are kept for
 * check_termination. */
void update_info(OSQPSolver* solver, OSQPInt iter) {
  OSQPInfo* info = solver->info;
  OSQPWorkspace* work = solver->work;
  const OSQPInt m = work->data->m, uns = unscaled_termination(solver);
  const OSQPVectorf *x = work->x, *y = work->y, *z = work->z;
  OSQPFloat pr = 0., pr_s = 0., du = 0., du_s = 0., ta = 0., tb = 0., tq = 0., tAty = 0., tPx = 0.;
  ObjTerms ot;
  info->iter = iter;
  /* compute_prim_res */
  if (m) {
    OSQPMatrix_Axpy(work->data->A, x, work->Ax, 1.0, 0.0);
    OSQPVectorf_minus(work->z_prev, work->Ax, z);
    osqp_amd_red_norm_inf(work->z_prev, &pr);
    if (uns) osqp_amd_red_scaled_norm_inf(work->scaling->Einv, work->z_prev, &pr_s);
  }
  /* compute_dual_res */
  OSQPVectorf_copy(work->x_prev, work->data->q);
  OSQPMatrix_Axpy(work->data->P, x, work->Px, 1.0, 0.0);
  OSQPVectorf_plus(work->x_prev, work->x_prev, work->Px);
  if (m) {
    OSQPMatrix_Atxpy(work->data->A, y, work->Aty, 1.0, 0.0);
    OSQPVectorf_plus(work->x_prev, work->x_prev, work->Aty);
  }
  osqp_amd_red_norm_inf(work->x_prev, &du);
  if (uns) osqp_amd_red_scaled_norm_inf(work->scaling->Dinv, work->x_prev, &du_s);
  /* compute_obj_val_dual_gap */
  obj_terms_queue(solver, x, y, work->Adelta_x, &ot);
  /* compute_prim_tol / compute_dual_tol */
  if (uns) {
    if (m) {
      osqp_amd_red_scaled_norm_inf(work->scaling->Einv, z, &ta);
      osqp_amd_red_scaled_norm_inf(work->scaling->Einv, work->Ax, &tb);
    }
    osqp_amd_red_scaled_norm_inf(work->scaling->Dinv, work->data->q, &tq);
    osqp_amd_red_scaled_norm_inf(work->scaling->Dinv, work->Aty, &tAty);
    osqp_amd_red_scaled_norm_inf(work->scaling->Dinv, work->Px, &tPx);
  } else {
    if (m) {
      osqp_amd_red_norm_inf(z, &ta);
      osqp_amd_red_norm_inf(work->Ax, &tb);
    }
    osqp_amd_red_norm_inf(work->data->q, &tq);
    osqp_amd_red_norm_inf(work->Aty, &tAty);
    osqp_amd_red_norm_inf(work->Px, &tPx);
  }
  osqp_amd_red_wait();
  if (m) {
    work->scaled_prim_res = pr;
    info->prim_res = uns ? pr_s : pr;
  } else {
    info->prim_res = 0.;
  }
  work->scaled_dual_res = du;
  info->dual_res = uns ? work->scaling->cinv * du_s : du;
  obj_terms_finish(solver, &ot, &info->obj_val, &info->dual_obj_val, &info->duality_gap);
  work->tol_prim_norm = c_max(ta, tb);
  {
    OSQPFloat r = c_max(tq, tAty);
    r = c_max(r, tPx);
    work->tol_dual_norm = uns ? r * work->scaling->cinv : r;
  }
  work->tol_valid = 1;
  info->primdual_int += c_absval(info->duality_gap);
  info->solve_time = osqp_amd_toc(work);
  info->rel_kkt_error = c_max(c_max(info->dual_res, info->prim_res), info->duality_gap);
  work->summary_printed = 0;
}

/* auxil.c:676-762 (polishing branch): residuals and objective of the
 * polished point; the duality-gap integral is not accumulated */
void update_info_polish(OSQPSolver* solver) {
  OSQPInfo* info = solver->info;
  OSQPWorkspace* work = solver->work;
  if (work->data->m == 0) work->pol_prim_res = 0.;
  else work->pol_prim_res = compute_prim_res(solver, work->pol_x, work->pol_z);
  work->pol_dual_res = compute_dual_res(solver, work->pol_x, work->pol_y);
  compute_obj_val_dual_gap(solver, work->pol_x, work->pol_y, &work->pol_obj_val, &work->pol_dual_obj_val,
                           &work->pol_duality_gap);
  info->polish_time = osqp_amd_toc(work);
  info->rel_kkt_error = c_max(c_max(work->pol_dual_res, work->pol_prim_res), work->pol_duality_gap);
  work->summary_printed = 0;
}

/* auxil.c:765-783 */
void reset_info(OSQPInfo* info) {
  info->solve_time = 0.0;
  info->polish_time = 0.0;
  update_status(info, OSQP_UNSOLVED);
  info->rho_updates = 0;
}

/* auxil.c:800-806 */
void update_status(OSQPInfo* info, OSQPInt status_val) {
  info->status_val = status_val;
  strcpy(info->status, OSQP_STATUS_MESSAGE[status_val]);
}

/* auxil.c:808-945 */
OSQPInt check_termination(OSQPSolver* solver, OSQPInt approximate) {
  OSQPInfo* info = solver->info;
  OSQPSettings* settings = solver->settings;
  OSQPWorkspace* work = solver->work;
  OSQPFloat eps_abs = settings->eps_abs, eps_rel = settings->eps_rel;
  OSQPFloat eps_prim_inf = settings->eps_prim_inf, eps_dual_inf = settings->eps_dual_inf;
  OSQPInt prim_ok = 0, dual_ok = 0, gap_ok = 0, prim_inf = 0, dual_inf = 0;

  if ((info->prim_res > OSQP_INFTY) || (info->dual_res > OSQP_INFTY)) {
    update_status(info, OSQP_NON_CVX);
    info->obj_val = OSQP_NAN;
    return 1;
  }
  if (approximate) {
    eps_abs *= 10;
    eps_rel *= 10;
    eps_prim_inf *= 10;
    eps_dual_inf *= 10;
  }
  if (work->data->m == 0) {
    prim_ok = 1;
  } else {
    if (info->prim_res < compute_prim_tol(solver, eps_abs, eps_rel)) prim_ok = 1;
    else prim_inf = is_primal_infeasible(solver, eps_prim_inf);
  }
  if (info->dual_res < compute_dual_tol(solver, eps_abs, eps_rel)) dual_ok = 1;
  else dual_inf = is_dual_infeasible(solver, eps_dual_inf);
  if (settings->check_dualgap) {
    OSQPFloat eps_gap = compute_duality_gap_tol(solver, eps_abs, eps_rel);
    if (unscaled_termination(solver)) {
      if (c_absval(info->duality_gap) < eps_gap) gap_ok = 1;
    } else {
      if (c_absval(work->scaled_dual_gap) < eps_gap) gap_ok = 1;
    }
  } else {
    gap_ok = 1;
  }
  if (prim_ok && dual_ok && gap_ok) {
    update_status(info, approximate ? OSQP_SOLVED_INACCURATE : OSQP_SOLVED);
    return 1;
  } else if (prim_inf) {
    update_status(info, approximate ? OSQP_PRIMAL_INFEASIBLE_INACCURATE : OSQP_PRIMAL_INFEASIBLE);
    if (unscaled_termination(solver)) OSQPVectorf_ew_prod(work->delta_y, work->delta_y, work->scaling->E);
    info->obj_val = OSQP_INFTY;
    return 1;
  } else if (dual_inf) {
    update_status(info, approximate ? OSQP_DUAL_INFEASIBLE_INACCURATE : OSQP_DUAL_INFEASIBLE);
    if (unscaled_termination(solver)) OSQPVectorf_ew_prod(work->delta_x, work->delta_x, work->scaling->D);
    info->obj_val = -OSQP_INFTY;
    return 1;
  }
  return 0;
}

/* auxil.c:950-1017 — reads the CSC arrays on the host */
OSQPInt validate_data(const OSQPCscMatrix* P, const OSQPFloat* q, const OSQPCscMatrix* A, const OSQPFloat* l,
                      const OSQPFloat* u, OSQPInt m, OSQPInt n) {
  OSQPInt j, ptr;
  if (!P) { fprintf(stderr, "Missing quadratic cost matrix P\n"); return 1; }
  if (!A) { fprintf(stderr, "Missing constraint matrix A\n"); return 1; }
  if (!q) { fprintf(stderr, "Missing linear cost vector q\n"); return 1; }
  if ((n <= 0) || (m < 0)) { fprintf(stderr, "n must be positive and m nonnegative; n = %i, m = %i\n", (int)n, (int)m); return 1; }
  if (P->m != n) { fprintf(stderr, "P does not have dimension n x n with n = %i\n", (int)n); return 1; }
  if (P->m != P->n) { fprintf(stderr, "P is not square\n"); return 1; }
  for (j = 0; j < n; j++)
    for (ptr = P->p[j]; ptr < P->p[j + 1]; ptr++)
      if (P->i[ptr] > j) { fprintf(stderr, "P is not upper triangular\n"); return 1; }
  if ((A->m != m) || (A->n != n)) { fprintf(stderr, "A does not have dimension %i x %i\n", (int)m, (int)n); return 1; }
  for (j = 0; j < m; j++)
    if (l[j] > u[j]) {
      fprintf(stderr, "Lower bound at index %d is greater than upper bound: %.4e > %.4e\n", (int)j, l[j], u[j]);
      return 1;
    }
  return 0;
}

/* auxil.c:1019-1037 */
static OSQPInt validate_linsys_solver(OSQPInt linsys_solver) {
  if ((linsys_solver == OSQP_INDIRECT_SOLVER) && (osqp_algebra_linsys_supported() & OSQP_CAPABILITY_INDIRECT_SOLVER))
    return 0;
  if ((linsys_solver == OSQP_DIRECT_SOLVER) && (osqp_algebra_linsys_supported() & OSQP_CAPABILITY_DIRECT_SOLVER))
    return 0;
  return 1;
}

#define REQUIRE(cond, msg) do { if (!(cond)) { fprintf(stderr, "%s\n", msg); return 1; } } while (0)

/* auxil.c:1040-1223 */
OSQPInt validate_settings(const OSQPSettings* s, OSQPInt from_setup) {
  REQUIRE(s, "Missing settings!");
  REQUIRE(!(from_setup && validate_linsys_solver(s->linsys_solver)), "linsys_solver not recognized");
  REQUIRE(!(from_setup && s->allocate_solution != 0 && s->allocate_solution != 1), "allocate_solution must be either 0 or 1");
  REQUIRE(s->verbose == 0 || s->verbose == 1, "verbose must be either 0 or 1");
  REQUIRE(s->profiler_level == 0 || s->profiler_level == 1 || s->profiler_level == 2, "profiler_level must be either 0, 1 or 2");
  REQUIRE(s->warm_starting == 0 || s->warm_starting == 1, "warm_start must be either 0 or 1");
  REQUIRE(!(from_setup && s->scaling < 0), "scaling must be nonnegative");
  REQUIRE(s->polishing == 0 || s->polishing == 1, "polishing must be either 0 or 1");
  REQUIRE(!(from_setup && s->rho <= 0.0), "rho must be positive");
  REQUIRE(!(from_setup && s->rho_is_vec != 0 && s->rho_is_vec != 1), "rho_is_vec must be either 0 or 1");
  REQUIRE(!(from_setup && s->sigma <= 0.0), "sigma must be positive");
  REQUIRE(!(s->alpha <= 0.0 || s->alpha >= 2.0), "alpha must be strictly between 0 and 2");
  REQUIRE(s->cg_max_iter > 0, "cg_max_iter must be positive");
  REQUIRE(s->cg_tol_reduction > 0, "cg_tol_reduction must be positive");
  REQUIRE(!(s->cg_tol_fraction <= 0.0 || s->cg_tol_fraction >= 1.0), "cg_tol_fraction must be strictly between 0 and 1");
  /* auxil.c:1130-1135 (the reference's operator precedence is kept) */
  REQUIRE(!((from_setup && s->adaptive_rho < 0) || s->adaptive_rho >= _OSQP_ADAPTIVE_RHO_UPDATE_LAST_VALUE),
          "adaptive_rho not a valid choice");
  REQUIRE(!(from_setup && s->adaptive_rho_interval < 0), "adaptive_rho_interval must be nonnegative");
  REQUIRE(!(from_setup && s->adaptive_rho_fraction <= 0), "adaptive_rho_fraction must be positive");
  REQUIRE(!(from_setup && s->adaptive_rho_tolerance < 1.0), "adaptive_rho_tolerance must be >= 1");
  REQUIRE(s->max_iter > 0, "max_iter must be positive");
  REQUIRE(s->eps_abs >= 0.0, "eps_abs must be nonnegative");
  REQUIRE(s->eps_rel >= 0.0, "eps_rel must be nonnegative");
  REQUIRE(!(s->eps_rel == 0.0 && s->eps_abs == 0.0), "at least one of eps_abs and eps_rel must be positive");
  REQUIRE(s->eps_prim_inf > 0.0, "eps_prim_inf must be positive");
  REQUIRE(s->eps_dual_inf > 0.0, "eps_dual_inf must be positive");
  REQUIRE(s->scaled_termination == 0 || s->scaled_termination == 1, "scaled_termination must be either 0 or 1");
  REQUIRE(s->check_termination >= 0, "check_termination must be nonnegative");
  REQUIRE(s->check_dualgap == 0 || s->check_dualgap == 1, "check_dualgap must be either 0 or 1");
  REQUIRE(s->time_limit > 0.0, "time_limit must be positive");
  REQUIRE(s->delta > 0.0, "delta must be positive");
  REQUIRE(s->polish_refine_iter >= 0, "polish_refine_iter must be nonnegative");
  return 0;
}

/* ---- Ruiz equilibration: src/scaling.c ------------------------------------ */

static OSQPFloat limit_scaling_scalar(OSQPFloat v) {
  v = v < OSQP_MIN_SCALING ? 1.0 : v;
  v = v > OSQP_MAX_SCALING ? OSQP_MAX_SCALING : v;
  return v;
}

static void limit_scaling_vector(OSQPVectorf* v) {
  OSQPVectorf_set_scalar_if_lt(v, v, OSQP_MIN_SCALING, 1.0);
  OSQPVectorf_set_scalar_if_gt(v, v, OSQP_MAX_SCALING, OSQP_MAX_SCALING);
}

/* scaling.c:30-46 */
static void compute_inf_norm_cols_KKT(const OSQPMatrix* P, const OSQPMatrix* A, OSQPVectorf* D,
                                      OSQPVectorf* D_temp_A, OSQPVectorf* E) {
  OSQPMatrix_col_norm_inf(P, D);
  OSQPMatrix_col_norm_inf(A, D_temp_A);
  OSQPVectorf_ew_max_vec(D, D_temp_A, D);
  OSQPMatrix_row_norm_inf(A, E);
}

/* scaling.c:48-164 */
OSQPInt scale_data(OSQPSolver* solver) {
  OSQPSettings* settings = solver->settings;
  OSQPWorkspace* work = solver->work;
  OSQPInt i, n = work->data->n;
  OSQPFloat c_temp, inf_norm_q;
  work->scaling->c = 1.0;
  OSQPVectorf_set_scalar(work->scaling->D, 1.);
  OSQPVectorf_set_scalar(work->scaling->Dinv, 1.);
  OSQPVectorf_set_scalar(work->scaling->E, 1.);
  OSQPVectorf_set_scalar(work->scaling->Einv, 1.);
  for (i = 0; i < settings->scaling; i++) {
    compute_inf_norm_cols_KKT(work->data->P, work->data->A, work->D_temp, work->D_temp_A, work->E_temp);
    limit_scaling_vector(work->D_temp);
    limit_scaling_vector(work->E_temp);
    OSQPVectorf_ew_sqrt(work->D_temp);
    OSQPVectorf_ew_sqrt(work->E_temp);
    OSQPVectorf_ew_reciprocal(work->D_temp, work->D_temp);
    OSQPVectorf_ew_reciprocal(work->E_temp, work->E_temp);
    OSQPMatrix_lmult_diag(work->data->P, work->D_temp);
    OSQPMatrix_rmult_diag(work->data->P, work->D_temp);
    OSQPMatrix_lmult_diag(work->data->A, work->E_temp);
    OSQPMatrix_rmult_diag(work->data->A, work->D_temp);
    OSQPVectorf_ew_prod(work->data->q, work->data->q, work->D_temp);
    OSQPVectorf_ew_prod(work->scaling->D, work->scaling->D, work->D_temp);
    OSQPVectorf_ew_prod(work->scaling->E, work->scaling->E, work->E_temp);
    OSQPMatrix_col_norm_inf(work->data->P, work->D_temp);
    c_temp = OSQPVectorf_norm_1(work->D_temp);
    c_temp = c_temp / n;
    inf_norm_q = OSQPVectorf_norm_inf(work->data->q);
    inf_norm_q = limit_scaling_scalar(inf_norm_q);
    c_temp = c_max(c_temp, inf_norm_q);
    c_temp = limit_scaling_scalar(c_temp);
    c_temp = 1. / c_temp;
    OSQPMatrix_mult_scalar(work->data->P, c_temp);
    OSQPVectorf_mult_scalar(work->data->q, c_temp);
    work->scaling->c *= c_temp;
  }
  work->scaling->cinv = 1. / work->scaling->c;
  OSQPVectorf_ew_reciprocal(work->scaling->Dinv, work->scaling->D);
  OSQPVectorf_ew_reciprocal(work->scaling->Einv, work->scaling->E);
  OSQPVectorf_ew_prod(work->data->l, work->data->l, work->scaling->E);
  OSQPVectorf_ew_prod(work->data->u, work->data->u, work->scaling->E);
  return 0;
}

/* scaling.c:169-192 */
OSQPInt unscale_data(OSQPSolver* solver) {
  OSQPWorkspace* work = solver->work;
  OSQPMatrix_mult_scalar(work->data->P, work->scaling->cinv);
  OSQPMatrix_lmult_diag(work->data->P, work->scaling->Dinv);
  OSQPMatrix_rmult_diag(work->data->P, work->scaling->Dinv);
  OSQPVectorf_mult_scalar(work->data->q, work->scaling->cinv);
  OSQPVectorf_ew_prod(work->data->q, work->data->q, work->scaling->Dinv);
  OSQPMatrix_lmult_diag(work->data->A, work->scaling->Einv);
  OSQPMatrix_rmult_diag(work->data->A, work->scaling->Dinv);
  OSQPVectorf_ew_prod(work->data->l, work->data->l, work->scaling->Einv);
  OSQPVectorf_ew_prod(work->data->u, work->data->u, work->scaling->Einv);
  return 0;
}

/* scaling.c:194-208 */
OSQPInt unscale_solution(OSQPVectorf* usolx, OSQPVectorf* usoly, const OSQPVectorf* solx, const OSQPVectorf* soly,
                         OSQPWorkspace* work) {
  OSQPVectorf_ew_prod(usolx, solx, work->scaling->D);
  OSQPVectorf_ew_prod(usoly, soly, work->scaling->E);
  OSQPVectorf_mult_scalar(usoly, work->scaling->cinv);
  return 0;
}
