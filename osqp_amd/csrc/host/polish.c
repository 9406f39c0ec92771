/*
 * polish.c — solution polishing (src/polish.c:19-515) on the GPU backend.
 *
 * The active set is guessed from the ADMM solution, the reduced KKT system
 * [P + delta I, Ared'; Ared, -delta I] is solved by the backend's linear
 * solver in polishing mode (the PCG on P + delta I + Ared' Ared / delta, FP64
 * residuals), the regularisation error is removed by iterative refinement, and
 * the polished point is kept when it lowers the residuals.
 */
#include <stdlib.h>

#include "host_internal.h"

/* polish.c:19-107 — active flags (-1 lower or equality, +1 upper, 0 free) */
static OSQPInt form_Ared(OSQPWorkspace* work) {
  OSQPInt j, m = work->data->m, n_active = 0;
  OSQPInt* flags = (OSQPInt*)malloc(sizeof(OSQPInt) * (size_t)(m > 0 ? m : 1));
  OSQPFloat* z = (OSQPFloat*)malloc(sizeof(OSQPFloat) * (size_t)(m > 0 ? m : 1));
  OSQPFloat* y = (OSQPFloat*)malloc(sizeof(OSQPFloat) * (size_t)(m > 0 ? m : 1));
  OSQPFloat* l = (OSQPFloat*)malloc(sizeof(OSQPFloat) * (size_t)(m > 0 ? m : 1));
  OSQPFloat* u = (OSQPFloat*)malloc(sizeof(OSQPFloat) * (size_t)(m > 0 ? m : 1));
  OSQPInt rc = OSQP_NO_ERROR;
  if (!flags || !z || !y || !l || !u) {
    rc = osqp_error(OSQP_MEM_ALLOC_ERROR);
    goto done;
  }
  OSQPVectorf_to_raw(z, work->z);
  OSQPVectorf_to_raw(y, work->y);
  OSQPVectorf_to_raw(l, work->data->l);
  OSQPVectorf_to_raw(u, work->data->u);
  for (j = 0; j < m; j++) {
    if ((z[j] - l[j] < -y[j]) || (l[j] == u[j])) {
      flags[j] = -1;
      n_active++;
    } else if (u[j] - z[j] < y[j]) {
      flags[j] = +1;
      n_active++;
    } else {
      flags[j] = 0;
    }
  }
  OSQPVectori_from_raw(work->pol_active, flags);
  work->pol_n_active = n_active;
  work->pol_Ared = OSQPMatrix_submatrix_byrows(work->data->A, work->pol_active);
  if (!work->pol_Ared) rc = osqp_error(OSQP_MEM_ALLOC_ERROR);
done:
  free(flags);
  free(z);
  free(y);
  free(l);
  free(u);
  return rc;
}

/* polish.c:115-175 — rhs_red = [-q; l of lower-active rows; u of upper-active rows] */
static OSQPInt form_rhs_red(OSQPWorkspace* work, OSQPVectorf* rhs) {
  OSQPInt j, counter = 0, n = work->data->n, m = work->data->m;
  OSQPInt len = OSQPVectorf_length(rhs);
  OSQPInt* flags = (OSQPInt*)malloc(sizeof(OSQPInt) * (size_t)(m > 0 ? m : 1));
  OSQPFloat* r = (OSQPFloat*)malloc(sizeof(OSQPFloat) * (size_t)(len > 0 ? len : 1));
  OSQPFloat* q = (OSQPFloat*)malloc(sizeof(OSQPFloat) * (size_t)(n > 0 ? n : 1));
  OSQPFloat* l = (OSQPFloat*)malloc(sizeof(OSQPFloat) * (size_t)(m > 0 ? m : 1));
  OSQPFloat* u = (OSQPFloat*)malloc(sizeof(OSQPFloat) * (size_t)(m > 0 ? m : 1));
  OSQPInt rc = OSQP_NO_ERROR;
  if (!flags || !r || !q || !l || !u) {
    rc = osqp_error(OSQP_MEM_ALLOC_ERROR);
    goto done;
  }
  OSQPVectori_to_raw(flags, work->pol_active);
  OSQPVectorf_to_raw(q, work->data->q);
  OSQPVectorf_to_raw(l, work->data->l);
  OSQPVectorf_to_raw(u, work->data->u);
  for (j = 0; j < n; j++) r[j] = -q[j];
  for (j = 0; j < m; j++) {
    if (flags[j] == -1) r[n + counter++] = l[j];
    else if (flags[j] == 1) r[n + counter++] = u[j];
  }
  OSQPVectorf_from_raw(rhs, r);
done:
  free(flags);
  free(r);
  free(q);
  free(l);
  free(u);
  return rc;
}

/* polish.c:187-250 — z += (K + dK)^{-1} (b - K z), polish_refine_iter times */
static OSQPInt iterative_refinement(OSQPSolver* solver, LinSysSolver* p, OSQPVectorf* z, OSQPVectorf* b) {
  OSQPWorkspace* work = solver->work;
  OSQPInt i, n = work->data->n, mred;
  OSQPVectorf *rhs, *rhs1, *rhs2, *z1, *z2;
  if (solver->settings->polish_refine_iter <= 0) return 0;
  mred = OSQPMatrix_get_m(work->pol_Ared);
  rhs = OSQPVectorf_malloc(n + mred);
  if (!rhs) return osqp_error(OSQP_MEM_ALLOC_ERROR);
  rhs1 = OSQPVectorf_view(rhs, 0, n);
  rhs2 = OSQPVectorf_view(rhs, n, mred);
  z1 = OSQPVectorf_view(z, 0, n);
  z2 = OSQPVectorf_view(z, n, mred);
  if (!rhs1 || !rhs2 || !z1 || !z2) {
    OSQPVectorf_free(rhs);
    OSQPVectorf_view_free(rhs1);
    OSQPVectorf_view_free(rhs2);
    OSQPVectorf_view_free(z1);
    OSQPVectorf_view_free(z2);
    return osqp_error(OSQP_MEM_ALLOC_ERROR);
  }
  for (i = 0; i < solver->settings->polish_refine_iter; i++) {
    OSQPVectorf_copy(rhs, b);
    OSQPMatrix_Axpy(work->data->P, z1, rhs1, -1.0, 1.0);
    OSQPMatrix_Atxpy(work->pol_Ared, z2, rhs1, -1.0, 1.0);
    OSQPMatrix_Axpy(work->pol_Ared, z1, rhs2, -1.0, 1.0);
    p->solve(p, rhs, 1);
    OSQPVectorf_plus(z, z, rhs);
  }
  OSQPVectorf_free(rhs);
  OSQPVectorf_view_free(rhs1);
  OSQPVectorf_view_free(rhs2);
  OSQPVectorf_view_free(z1);
  OSQPVectorf_view_free(z2);
  return 0;
}

/* polish.c:258-311 — y = 0 on free rows, y_red on the active ones */
static OSQPInt get_ypol_from_yred(OSQPWorkspace* work, OSQPVectorf* yred_vf) {
  OSQPInt j, counter = 0, m = work->data->m, mred = OSQPVectorf_length(yred_vf);
  OSQPInt* flags;
  OSQPFloat *y, *yred;
  if (work->pol_n_active == 0) {
    OSQPVectorf_set_scalar(work->pol_y, 0.);
    return OSQP_NO_ERROR;
  }
  flags = (OSQPInt*)malloc(sizeof(OSQPInt) * (size_t)(m > 0 ? m : 1));
  y = (OSQPFloat*)malloc(sizeof(OSQPFloat) * (size_t)(m > 0 ? m : 1));
  yred = (OSQPFloat*)malloc(sizeof(OSQPFloat) * (size_t)(mred > 0 ? mred : 1));
  if (!flags || !y || !yred) {
    free(flags);
    free(y);
    free(yred);
    return osqp_error(OSQP_MEM_ALLOC_ERROR);
  }
  OSQPVectori_to_raw(flags, work->pol_active);
  OSQPVectorf_to_raw(yred, yred_vf);
  for (j = 0; j < m; j++) y[j] = flags[j] == 0 ? 0.0 : yred[counter++];
  OSQPVectorf_from_raw(work->pol_y, y);
  free(flags);
  free(y);
  free(yred);
  return OSQP_NO_ERROR;
}

/* polish.c:313-515 */
OSQPInt polish(OSQPSolver* solver) {
  OSQPInfo* info = solver->info;
  OSQPWorkspace* work = solver->work;
  OSQPInt n = work->data->n, exitflag, polish_successful;
  LinSysSolver* plsh = NULL;
  OSQPVectorf *rhs_red = NULL, *pol_sol = NULL, *xv = NULL, *yv = NULL;
  struct timespec t0 = work->tic;

  osqp_amd_tic(work);
  exitflag = form_Ared(work);
  if (exitflag) {
    info->status_polish = OSQP_POLISH_FAILED;
    work->tic = t0;
    return exitflag;
  }
  if (work->pol_n_active == 0) {
    info->status_polish = OSQP_POLISH_NO_ACTIVE_SET_FOUND;
    OSQPMatrix_free(work->pol_Ared);
    work->pol_Ared = NULL;
    work->tic = t0;
    return OSQP_NO_ERROR;
  }
  exitflag = osqp_algebra_init_linsys_solver(&plsh, work->data->P, work->pol_Ared, NULL, solver->settings, NULL,
                                             NULL, 1);
  if (exitflag) {
    info->status_polish = OSQP_POLISH_LINSYS_ERROR;
    OSQPMatrix_free(work->pol_Ared);
    work->pol_Ared = NULL;
    work->tic = t0;
    return exitflag;
  }
  rhs_red = OSQPVectorf_malloc(n + work->pol_n_active);
  exitflag = rhs_red ? form_rhs_red(work, rhs_red) : osqp_error(OSQP_MEM_ALLOC_ERROR);
  if (!exitflag) {
    pol_sol = OSQPVectorf_copy_new(rhs_red);
    if (pol_sol) {
      xv = OSQPVectorf_view(pol_sol, 0, n);
      yv = OSQPVectorf_view(pol_sol, n, work->pol_n_active);
    }
    if (!pol_sol || !xv || !yv) exitflag = osqp_error(OSQP_MEM_ALLOC_ERROR);
  }
  if (!exitflag) {
    plsh->warm_start(plsh, work->x);
    plsh->solve(plsh, pol_sol, 1);
    exitflag = iterative_refinement(solver, plsh, pol_sol, rhs_red);
  }
  if (exitflag) {
    info->status_polish = OSQP_POLISH_FAILED;
  } else {
    OSQPVectorf_copy(work->pol_x, xv);
    OSQPMatrix_Axpy(work->data->A, work->pol_x, work->pol_z, 1.0, 0.0);
    get_ypol_from_yred(work, yv);
    /* z in C and y in the normal cone of C at z: y += z; z = proj_C(y); y -= z */
    OSQPVectorf_plus(work->pol_y, work->pol_y, work->pol_z);
    OSQPVectorf_ew_bound_vec(work->pol_z, work->pol_y, work->data->l, work->data->u);
    OSQPVectorf_minus(work->pol_y, work->pol_y, work->pol_z);
    update_info_polish(solver);
    polish_successful = (work->pol_prim_res < info->prim_res && work->pol_dual_res < info->dual_res) ||
                        (work->pol_prim_res < info->prim_res && info->dual_res < 1e-10) ||
                        (work->pol_dual_res < info->dual_res && info->prim_res < 1e-10);
    if (polish_successful) {
      info->obj_val = work->pol_obj_val;
      info->dual_obj_val = work->pol_dual_obj_val;
      info->duality_gap = work->pol_duality_gap;
      info->prim_res = work->pol_prim_res;
      info->dual_res = work->pol_dual_res;
      info->status_polish = OSQP_POLISH_SUCCESS;
      OSQPVectorf_copy(work->x, work->pol_x);
      OSQPVectorf_copy(work->z, work->pol_z);
      OSQPVectorf_copy(work->y, work->pol_y);
    } else {
      info->status_polish = OSQP_POLISH_FAILED;
    }
  }
  if (plsh) plsh->free(plsh);
  OSQPMatrix_free(work->pol_Ared);
  work->pol_Ared = NULL;
  OSQPVectorf_free(rhs_red);
  OSQPVectorf_view_free(xv);
  OSQPVectorf_view_free(yv);
  OSQPVectorf_free(pol_sol);
  work->tic = t0;
  return exitflag > 0 ? exitflag : OSQP_NO_ERROR;
}
