/*
 * osqp_api.c — public entry points of the host core (include/osqp_amd.h).
 *
 * Restates src/osqp_api.c (setup 332-631, solve 636-1038, cleanup 1057-1174,
 * updates 1184-1530) on top of the algebra/hip backend.  The ADMM loop issues
 * its per-iteration vector updates and the PCG solve asynchronously on the
 * backend stream; the host waits for the GPU only where the algorithm needs a
 * scalar (termination checks, rho adaptation, infeasibility tests).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"
#include "../profilers.h"

OSQPInt osqp_capabilities(void) {
  return osqp_algebra_linsys_supported() | OSQP_CAPABILITY_UPDATE_MATRICES;
}

const char* osqp_version(void) { return "1.0.0-amd"; }

const char* osqp_error_message(OSQPInt error_flag) {
  if (error_flag <= 0 || error_flag >= OSQP_LAST_ERROR_PLACE) return "Unknown error code.";
  return OSQP_ERROR_MESSAGE[error_flag];
}

void osqp_get_dimensions(OSQPSolver* solver, OSQPInt* m, OSQPInt* n) {
  if (!solver || !solver->work || !solver->work->data) {
    *m = -1;
    *n = -1;
  } else {
    *m = solver->work->data->m;
    *n = solver->work->data->n;
  }
}

/* src/osqp_api.c:283-327 with this backend's defaults (indirect solver) */
void osqp_set_default_settings(OSQPSettings* settings) {
  if (!settings) return;
  settings->device = 0;
  settings->linsys_solver = osqp_algebra_default_linsys();
  settings->allocate_solution = 1;
  settings->profiler_level = 0;
  settings->verbose = OSQP_VERBOSE;
  settings->warm_starting = OSQP_WARM_STARTING;
  settings->scaling = OSQP_SCALING;
  settings->polishing = OSQP_POLISHING;
  settings->rho = (OSQPFloat)OSQP_RHO;
  settings->rho_is_vec = OSQP_RHO_IS_VEC;
  settings->sigma = (OSQPFloat)OSQP_SIGMA;
  settings->alpha = (OSQPFloat)OSQP_ALPHA;
  settings->cg_max_iter = OSQP_CG_MAX_ITER;
  settings->cg_tol_reduction = OSQP_CG_TOL_REDUCTION;
  settings->cg_tol_fraction = OSQP_CG_TOL_FRACTION;
  settings->cg_precond = OSQP_DIAGONAL_PRECONDITIONER;
  settings->adaptive_rho = OSQP_ADAPTIVE_RHO_UPDATE_DEFAULT;
  settings->adaptive_rho_interval = OSQP_ADAPTIVE_RHO_INTERVAL;
  settings->adaptive_rho_fraction = (OSQPFloat)OSQP_ADAPTIVE_RHO_FRACTION;
  settings->adaptive_rho_tolerance = (OSQPFloat)OSQP_ADAPTIVE_RHO_TOLERANCE;
  settings->max_iter = OSQP_MAX_ITER;
  settings->eps_abs = (OSQPFloat)OSQP_EPS_ABS;
  settings->eps_rel = (OSQPFloat)OSQP_EPS_REL;
  settings->eps_prim_inf = (OSQPFloat)OSQP_EPS_PRIM_INF;
  settings->eps_dual_inf = (OSQPFloat)OSQP_EPS_DUAL_INF;
  settings->scaled_termination = OSQP_SCALED_TERMINATION;
  settings->check_termination = OSQP_CHECK_TERMINATION;
  settings->check_dualgap = OSQP_CHECK_DUALGAP;
  settings->time_limit = OSQP_TIME_LIMIT;
  settings->delta = OSQP_DELTA;
  settings->polish_refine_iter = OSQP_POLISH_REFINE_ITER;
}

OSQPSettings* OSQPSettings_new(void) {
  OSQPSettings* s = (OSQPSettings*)calloc(1, sizeof(OSQPSettings));
  if (s) osqp_set_default_settings(s);
  return s;
}

void OSQPSettings_free(OSQPSettings* settings) { free(settings); }

void OSQPCscMatrix_set_data(OSQPCscMatrix* M, OSQPInt m, OSQPInt n, OSQPInt nzmax, OSQPFloat* x, OSQPInt* i,
                            OSQPInt* p) {
  M->m = m;
  M->n = n;
  M->nz = -1;
  M->nzmax = nzmax;
  M->x = x;
  M->i = i;
  M->p = p;
  M->owned = 0;
}

OSQPCscMatrix* OSQPCscMatrix_new(OSQPInt m, OSQPInt n, OSQPInt nzmax, OSQPFloat* x, OSQPInt* i, OSQPInt* p) {
  OSQPCscMatrix* M = (OSQPCscMatrix*)calloc(1, sizeof(OSQPCscMatrix));
  if (M) OSQPCscMatrix_set_data(M, m, n, nzmax, x, i, p);
  return M;
}

void OSQPCscMatrix_free(OSQPCscMatrix* M) {
  if (!M) return;
  if (M->owned) {
    free(M->p);
    free(M->i);
    free(M->x);
  }
  free(M);
}

static OSQPSettings* copy_settings(const OSQPSettings* s) {
  OSQPSettings* c = (OSQPSettings*)malloc(sizeof(OSQPSettings));
  if (c) memcpy(c, s, sizeof(OSQPSettings));
  return c;
}

/* ---- setup: src/osqp_api.c:332-631 ------------------------------------------ */
#define ALLOC_OR_FAIL(ptr) do { if (!(ptr)) return osqp_error(OSQP_MEM_ALLOC_ERROR); } while (0)

static OSQPInt setup_workspace(OSQPSolver** solverp, const OSQPCscMatrix* P, const OSQPFloat* q,
                               const OSQPCscMatrix* A, const OSQPFloat* l, const OSQPFloat* u, OSQPInt m, OSQPInt n,
                               const OSQPSettings* settings);

/* Host copies of a problem's arrays for a caller that passes device memory
 * (an extension: the reference's setup reads host arrays, src/auxil.c:
 * 950-1017).  Host arrays are used in place. */
typedef struct {
  OSQPCscMatrix P, A;
  const OSQPFloat *q, *l, *u;
  void* owned[9];
  int n_owned;
} host_inputs;

static void* host_copy(host_inputs* H, const void* src, size_t bytes, int* fail) {
  void* dst = malloc(bytes ? bytes : 1);
  if (!dst) {
    *fail = 1;
    return NULL;
  }
  H->owned[H->n_owned++] = dst;
  if (bytes && osqp_amd_copy_to_host(dst, src, bytes)) *fail = 1;
  return dst;
}

static int make_host_inputs(host_inputs* H, const OSQPCscMatrix* P, const OSQPFloat* q, const OSQPCscMatrix* A,
                            const OSQPFloat* l, const OSQPFloat* u, OSQPInt m, OSQPInt n) {
  const void* ptrs[9];
  int k, dev = 0, fail = 0;
  memset(H, 0, sizeof(*H));
  if (!P || !A) return 0;
  H->P = *P;
  H->A = *A;
  H->q = q;
  H->l = l;
  H->u = u;
  ptrs[0] = P->p, ptrs[1] = P->i, ptrs[2] = P->x, ptrs[3] = A->p, ptrs[4] = A->i, ptrs[5] = A->x;
  ptrs[6] = q, ptrs[7] = l, ptrs[8] = u;
  for (k = 0; k < 9; k++) dev |= osqp_amd_ptr_is_device(ptrs[k]) ? 1 : 0;
  if (!dev || n < 0 || m < 0 || !P->p || !A->p) return 0;
  H->P.p = (OSQPInt*)host_copy(H, P->p, sizeof(OSQPInt) * ((size_t)n + 1), &fail);
  H->A.p = (OSQPInt*)host_copy(H, A->p, sizeof(OSQPInt) * ((size_t)n + 1), &fail);
  if (fail) return 1;
  H->P.i = (OSQPInt*)host_copy(H, P->i, sizeof(OSQPInt) * (size_t)H->P.p[n], &fail);
  H->P.x = (OSQPFloat*)host_copy(H, P->x, sizeof(OSQPFloat) * (size_t)H->P.p[n], &fail);
  H->A.i = (OSQPInt*)host_copy(H, A->i, sizeof(OSQPInt) * (size_t)H->A.p[n], &fail);
  H->A.x = (OSQPFloat*)host_copy(H, A->x, sizeof(OSQPFloat) * (size_t)H->A.p[n], &fail);
  if (q) H->q = (const OSQPFloat*)host_copy(H, q, sizeof(OSQPFloat) * (size_t)n, &fail);
  if (l) H->l = (const OSQPFloat*)host_copy(H, l, sizeof(OSQPFloat) * (size_t)m, &fail);
  if (u) H->u = (const OSQPFloat*)host_copy(H, u, sizeof(OSQPFloat) * (size_t)m, &fail);
  return fail;
}

static void free_host_inputs(host_inputs* H) {
  int k;
  for (k = 0; k < H->n_owned; k++) free(H->owned[k]);
  H->n_owned = 0;
}

OSQPInt osqp_setup(OSQPSolver** solverp, const OSQPCscMatrix* P, const OSQPFloat* q, const OSQPCscMatrix* A,
                   const OSQPFloat* l, const OSQPFloat* u, OSQPInt m, OSQPInt n, const OSQPSettings* settings) {
  OSQPInt exitflag;
  host_inputs H;
  if (make_host_inputs(&H, P, q, A, l, u, m, n)) {
    free_host_inputs(&H);
    return osqp_error(OSQP_DATA_VALIDATION_ERROR);
  }
  if (H.n_owned) {
    P = &H.P;
    A = &H.A;
    q = H.q;
    l = H.l;
    u = H.u;
  }
  exitflag = validate_data(P, q, A, l, u, m, n);
  if (exitflag) {
    free_host_inputs(&H);
    return osqp_error(OSQP_DATA_VALIDATION_ERROR);
  }
  if (validate_settings(settings, 1)) {
    free_host_inputs(&H);
    return osqp_error(OSQP_SETTINGS_VALIDATION_ERROR);
  }
  /* src/osqp_api.c:353-354,621 */
  osqp_amd_profiler_level(settings->profiler_level);
  osqp_amd_sec_push(OSQP_AMD_SEC_SETUP);
  exitflag = setup_workspace(solverp, P, q, A, l, u, m, n, settings);
  osqp_amd_sec_pop(OSQP_AMD_SEC_SETUP);
  free_host_inputs(&H);
  return exitflag;
}

static OSQPInt setup_workspace(OSQPSolver** solverp, const OSQPCscMatrix* P, const OSQPFloat* q,
                               const OSQPCscMatrix* A, const OSQPFloat* l, const OSQPFloat* u, OSQPInt m, OSQPInt n,
                               const OSQPSettings* settings) {
  OSQPInt exitflag;
  OSQPSolver* solver;
  OSQPWorkspace* work;

  solver = (OSQPSolver*)calloc(1, sizeof(OSQPSolver));
  ALLOC_OR_FAIL(solver);
  *solverp = solver;
  work = (OSQPWorkspace*)calloc(1, sizeof(OSQPWorkspace));
  ALLOC_OR_FAIL(work);
  solver->work = work;
  solver->info = (OSQPInfo*)calloc(1, sizeof(OSQPInfo));
  ALLOC_OR_FAIL(solver->info);
  osqp_amd_tic(work);

  exitflag = osqp_algebra_init_libs(settings->device);
  if (exitflag) {
    /* the workspace holds no device data yet: release it before reporting */
    free(solver->info);
    free(work);
    free(solver);
    *solverp = NULL;
    return osqp_error(OSQP_ALGEBRA_LOAD_ERROR);
  }

  work->data = (OSQPData*)calloc(1, sizeof(OSQPData));
  ALLOC_OR_FAIL(work->data);
  work->data->m = m;
  work->data->n = n;
  work->data->P = OSQPMatrix_new_from_csc(P, 1);
  work->data->q = OSQPVectorf_new(q, n);
  ALLOC_OR_FAIL(work->data->P && work->data->q);
  work->data->A = OSQPMatrix_new_from_csc(A, 0);
  ALLOC_OR_FAIL(work->data->A);
  work->data->l = OSQPVectorf_new(l, m);
  work->data->u = OSQPVectorf_new(u, m);
  ALLOC_OR_FAIL(work->data->l && work->data->u);

  if (settings->rho_is_vec) {
    work->rho_vec = OSQPVectorf_malloc(m);
    work->rho_inv_vec = OSQPVectorf_malloc(m);
    ALLOC_OR_FAIL(work->rho_vec && work->rho_inv_vec);
    work->constr_type = OSQPVectori_calloc(m);
    ALLOC_OR_FAIL(work->constr_type);
  } else {
    work->rho_vec = NULL;
    work->rho_inv_vec = NULL;
  }

  work->x = OSQPVectorf_calloc(n);
  work->z = OSQPVectorf_calloc(m);
  work->xz_tilde = OSQPVectorf_calloc(n + m);
  work->xtilde_view = OSQPVectorf_view(work->xz_tilde, 0, n);
  work->ztilde_view = OSQPVectorf_view(work->xz_tilde, n, m);
  work->x_prev = OSQPVectorf_calloc(n);
  work->z_prev = OSQPVectorf_calloc(m);
  work->y = OSQPVectorf_calloc(m);
  ALLOC_OR_FAIL(work->x && work->z && work->xz_tilde && work->xtilde_view && work->ztilde_view);
  ALLOC_OR_FAIL(work->x_prev && work->z_prev && work->y);
  work->Ax = OSQPVectorf_calloc(m);
  work->Px = OSQPVectorf_calloc(n);
  work->Aty = OSQPVectorf_calloc(n);
  work->delta_y = OSQPVectorf_calloc(m);
  work->Atdelta_y = OSQPVectorf_calloc(n);
  work->delta_x = OSQPVectorf_calloc(n);
  work->pol_active = OSQPVectori_calloc(m);
  work->pol_x = OSQPVectorf_calloc(n);
  work->pol_z = OSQPVectorf_calloc(m);
  work->pol_y = OSQPVectorf_calloc(m);
  ALLOC_OR_FAIL(work->pol_active && work->pol_x && work->pol_z && work->pol_y);
  work->Pdelta_x = OSQPVectorf_calloc(n);
  work->Adelta_x = OSQPVectorf_calloc(m);
  ALLOC_OR_FAIL(work->Ax && work->Px && work->Aty && work->delta_y && work->Atdelta_y);
  ALLOC_OR_FAIL(work->delta_x && work->Pdelta_x && work->Adelta_x);

  solver->settings = copy_settings(settings);
  ALLOC_OR_FAIL(solver->settings);

  if (settings->scaling) {
    work->scaling = (OSQPScaling*)calloc(1, sizeof(OSQPScaling));
    ALLOC_OR_FAIL(work->scaling);
    work->scaling->D = OSQPVectorf_calloc(n);
    work->scaling->Dinv = OSQPVectorf_calloc(n);
    work->scaling->E = OSQPVectorf_calloc(m);
    work->scaling->Einv = OSQPVectorf_calloc(m);
    ALLOC_OR_FAIL(work->scaling->D && work->scaling->Dinv && work->scaling->E && work->scaling->Einv);
    work->D_temp = OSQPVectorf_calloc(n);
    work->D_temp_A = OSQPVectorf_calloc(n);
    work->E_temp = OSQPVectorf_calloc(m);
    ALLOC_OR_FAIL(work->D_temp && work->D_temp_A && work->E_temp);
    osqp_amd_sec_push(OSQP_AMD_SEC_SCALE);
    scale_data(solver);
    osqp_amd_sec_pop(OSQP_AMD_SEC_SCALE);
  } else {
    work->scaling = NULL;
    work->D_temp = NULL;
    work->D_temp_A = NULL;
    work->E_temp = NULL;
  }

  if (settings->rho_is_vec) {
    set_rho_vec(solver);
  } else {
    solver->settings->rho = c_min(c_max(settings->rho, OSQP_RHO_MIN), OSQP_RHO_MAX);
    work->rho_inv = 1. / settings->rho;
  }

  osqp_amd_sec_push(OSQP_AMD_SEC_LINSYS_INIT);
  exitflag = osqp_algebra_init_linsys_solver(&(work->linsys_solver), work->data->P, work->data->A, work->rho_vec,
                                             solver->settings, &work->scaled_prim_res, &work->scaled_dual_res, 0);
  osqp_amd_sec_pop(OSQP_AMD_SEC_LINSYS_INIT);
  if (exitflag == OSQP_NONCVX_ERROR) {
    update_status(solver->info, OSQP_NON_CVX);
    return osqp_error(exitflag);
  } else if (exitflag) {
    return osqp_error(exitflag);
  }

  osqp_cold_start(solver);

  if (settings->allocate_solution) {
    solver->solution = (OSQPSolution*)calloc(1, sizeof(OSQPSolution));
    ALLOC_OR_FAIL(solver->solution);
    solver->solution->x = (OSQPFloat*)calloc(1, (size_t)(n > 0 ? n : 1) * sizeof(OSQPFloat));
    solver->solution->y = (OSQPFloat*)calloc(1, (size_t)(m > 0 ? m : 1) * sizeof(OSQPFloat));
    solver->solution->prim_inf_cert = (OSQPFloat*)calloc(1, (size_t)(m > 0 ? m : 1) * sizeof(OSQPFloat));
    solver->solution->dual_inf_cert = (OSQPFloat*)calloc(1, (size_t)(n > 0 ? n : 1) * sizeof(OSQPFloat));
    ALLOC_OR_FAIL(solver->solution->x && solver->solution->dual_inf_cert);
    ALLOC_OR_FAIL(solver->solution->y && solver->solution->prim_inf_cert);
  } else {
    solver->solution = NULL;
  }

  solver->info->status_polish = OSQP_POLISH_NOT_PERFORMED;
  update_status(solver->info, OSQP_UNSOLVED);
  solver->info->solve_time = 0.0;
  solver->info->update_time = 0.0;
  solver->info->polish_time = 0.0;
  solver->info->run_time = 0.0;
  osqp_amd_algebra_sync();
  solver->info->setup_time = osqp_amd_toc(work);
  work->first_run = 1;
  work->clear_update_time = 0;
  work->rho_update_from_solve = 0;
  solver->info->rho_updates = 0;
  solver->info->rho_estimate = solver->settings->rho;
  solver->info->obj_val = OSQP_INFTY;
  solver->info->prim_res = OSQP_INFTY;
  solver->info->dual_res = OSQP_INFTY;
  solver->info->rel_kkt_error = OSQP_INFTY;
  work->last_rel_kkt = OSQP_INFTY;
  work->rho_updated = 0;

  switch (solver->settings->adaptive_rho) {
    case OSQP_ADAPTIVE_RHO_UPDATE_ITERATIONS:
      if (solver->settings->adaptive_rho_interval == 0)
        solver->settings->adaptive_rho_interval = solver->settings->check_termination
                                                      ? OSQP_ADAPTIVE_RHO_MULTIPLE_TERMINATION *
                                                            solver->settings->check_termination
                                                      : OSQP_ADAPTIVE_RHO_FIXED;
      break;
    case OSQP_ADAPTIVE_RHO_UPDATE_KKT_ERROR:
      if (solver->settings->adaptive_rho_interval == 0) solver->settings->adaptive_rho_interval = 1;
      break;
    default:
      break;
  }
  work->summary_printed = 0;
  if (solver->settings->verbose) {
    char name[256];
    osqp_algebra_device_name(name, sizeof(name));
    printf("OSQP (MI355X HIP backend) n = %d, m = %d, device: %s\n", (int)n, (int)m, name);
  }
  return 0;
}

static void print_summary(const OSQPSolver* solver) {
  const OSQPInfo* info = solver->info;
  printf("%4d  % 12.4e  %9.2e  %9.2e  %9.2e  %.1es\n", (int)info->iter, info->obj_val, info->prim_res,
         info->dual_res, solver->settings->rho, info->solve_time);
  solver->work->summary_printed = 1;
}

/* ---- solve: src/osqp_api.c:636-1038 ---------------------------------------- */
OSQPInt osqp_solve(OSQPSolver* solver) {
  OSQPInt exitflag = 0, iter, can_check = 0, can_adapt = 0, can_print = 0;
  OSQPWorkspace* work;
  OSQPSettings* settings;
  OSQPFloat temp_run_time;

  if (!solver || !solver->work) return osqp_error(OSQP_WORKSPACE_NOT_INIT_ERROR);
  work = solver->work;
  settings = solver->settings;
  if (work->clear_update_time == 1) solver->info->update_time = 0.0;
  work->rho_update_from_solve = 1;
  osqp_amd_tic(work);
  osqp_amd_sec_push(OSQP_AMD_SEC_OPT_SOLVE);
  if (settings->verbose) printf("iter   objective    prim res   dual res   rho        time\n");
  osqp_amd_interrupt_listen();  /* src/osqp_api.c:690-693 */
  if (!settings->warm_starting) osqp_cold_start(solver);
  work->rhs_ready = 0;

  for (iter = 1; iter <= settings->max_iter; iter++) {
    osqp_amd_sec_push(OSQP_AMD_SEC_ADMM_ITER);
    swap_vectors(&(work->x), &(work->x_prev));
    swap_vectors(&(work->z), &(work->z_prev));
    work->tol_valid = 0;  /* the iterates change: update_info recomputes the tolerance norms */
    osqp_amd_sec_push(OSQP_AMD_SEC_ADMM_KKT_SOLVE);
    update_xz_tilde(solver, iter);
    osqp_amd_sec_pop(OSQP_AMD_SEC_ADMM_KKT_SOLVE);
    osqp_amd_sec_push(OSQP_AMD_SEC_ADMM_UPDATE);
    if (settings->rho_is_vec) {
      /* one fused kernel: the projection section covers the whole update */
      osqp_amd_sec_push(OSQP_AMD_SEC_ADMM_PROJ);
      update_xzy(solver);
      osqp_amd_sec_pop(OSQP_AMD_SEC_ADMM_PROJ);
    } else {
      update_x(solver);
      osqp_amd_sec_push(OSQP_AMD_SEC_ADMM_PROJ);
      update_z(solver);
      osqp_amd_sec_pop(OSQP_AMD_SEC_ADMM_PROJ);
      update_y(solver);
    }
    osqp_amd_sec_pop(OSQP_AMD_SEC_ADMM_UPDATE);
    osqp_amd_sec_pop(OSQP_AMD_SEC_ADMM_ITER);

    /* src/osqp_api.c:728-737 */
    if (osqp_amd_interrupted()) {
      update_status(solver->info, OSQP_SIGINT);
      if (settings->verbose) printf("Solver interrupted\n");
      exitflag = 1;
      goto exit;
    }

    if (work->first_run) temp_run_time = solver->info->setup_time + osqp_amd_toc(work);
    else temp_run_time = solver->info->update_time + osqp_amd_toc(work);
    if (settings->time_limit && (temp_run_time >= settings->time_limit)) {
      update_status(solver->info, OSQP_TIME_LIMIT_REACHED);
      can_print = 0;
      break;
    }

    can_check = settings->check_termination && (iter % settings->check_termination == 0);
    can_print = settings->verbose && ((iter % OSQP_PRINT_INTERVAL == 0) || (iter == 1));

    switch (settings->adaptive_rho) {
      case OSQP_ADAPTIVE_RHO_UPDATE_DISABLED:
        can_adapt = 0;
        break;
      case OSQP_ADAPTIVE_RHO_UPDATE_TIME:
        if (settings->adaptive_rho_interval == 0) {
          if (osqp_amd_toc(work) > settings->adaptive_rho_fraction * solver->info->setup_time) {
            OSQPInt base = settings->check_termination ? settings->check_termination : OSQP_CHECK_TERMINATION;
            OSQPFloat x = (OSQPFloat)iter + .5 * base;
            settings->adaptive_rho_interval = (OSQPInt)(x - fmod(x, (OSQPFloat)base));
            /* src/osqp_api.c:784-785 */
            settings->adaptive_rho_interval =
                c_min(1, c_max(settings->adaptive_rho_interval, settings->check_termination));
          } else {
            can_adapt = 0;
            break;
          }
        }
        /* fall through */
      case OSQP_ADAPTIVE_RHO_UPDATE_KKT_ERROR:
      case OSQP_ADAPTIVE_RHO_UPDATE_ITERATIONS:
        can_adapt = (settings->adaptive_rho_interval && (iter % settings->adaptive_rho_interval == 0)) ? 1 : 0;
        break;
    }

    if (can_check || can_print || can_adapt || iter == 1) update_info(solver, iter);
    if (can_check && check_termination(solver, 0)) break;

    work->rho_updated = 0;
    if (can_adapt && (settings->adaptive_rho == OSQP_ADAPTIVE_RHO_UPDATE_KKT_ERROR))
      can_adapt = (solver->info->rel_kkt_error <= settings->adaptive_rho_fraction * work->last_rel_kkt) ? 1 : 0;
    if (can_adapt) {
      osqp_amd_event_mark(OSQP_AMD_EVENT_RHO_UPDATE);
      if (adapt_rho(solver)) {
        fprintf(stderr, "ERROR in osqp_solve: Failed rho update\n");
        exitflag = 1;
        goto exit;
      }
    }
    if (work->rho_updated) work->last_rel_kkt = solver->info->rel_kkt_error;
    if (can_print || (settings->verbose && work->rho_updated)) print_summary(solver);
  }

  if (!can_check) {
    if (!can_print) update_info(solver, iter - 1);
    if (settings->verbose && !work->summary_printed) print_summary(solver);
    check_termination(solver, 0);
  }

  if (has_solution(solver->info))
    compute_obj_val_dual_gap(solver, work->x, work->y, &(solver->info->obj_val), &(solver->info->dual_obj_val),
                             &(solver->info->duality_gap));

  if (settings->verbose && !work->summary_printed) print_summary(solver);

  if (solver->info->status_val == OSQP_UNSOLVED) {
    if (!check_termination(solver, 1)) update_status(solver->info, OSQP_MAX_ITER_REACHED);
  }
  if (solver->info->status_val == OSQP_TIME_LIMIT_REACHED) {
    if (!check_termination(solver, 1)) update_status(solver->info, OSQP_TIME_LIMIT_REACHED);
  }

  solver->info->rho_estimate = compute_rho_estimate(solver);
  osqp_amd_algebra_sync();
  solver->info->solve_time = osqp_amd_toc(work);

  /* src/osqp_api.c:981-990 */
  solver->info->polish_time = 0.0;
  if (settings->polishing && (solver->info->status_val == OSQP_SOLVED)) {
    osqp_amd_sec_push(OSQP_AMD_SEC_POLISH);
    exitflag = polish(solver);
    osqp_amd_sec_pop(OSQP_AMD_SEC_POLISH);
    if (exitflag > 0) {
      fprintf(stderr, "ERROR in osqp_solve: Failed polishing\n");
      goto exit;
    }
  }

  if (work->first_run)
    solver->info->run_time = solver->info->setup_time + solver->info->solve_time + solver->info->polish_time;
  else
    solver->info->run_time = solver->info->update_time + solver->info->solve_time + solver->info->polish_time;
  if (work->first_run) work->first_run = 0;
  work->clear_update_time = 1;
  work->rho_update_from_solve = 0;
  if (settings->verbose) printf("status: %s, iterations: %d, objective: %.6e\n", solver->info->status,
                                (int)solver->info->iter, solver->info->obj_val);
  store_solution(solver, solver->solution);

exit:
  osqp_amd_interrupt_restore();
  osqp_amd_sec_pop(OSQP_AMD_SEC_OPT_SOLVE);
  return exitflag;
}

/* src/osqp_api.c:1041-1054 */
OSQPInt osqp_get_solution(OSQPSolver* solver, OSQPSolution* solution) {
  if (!solver || !solver->work || !solver->settings || !solver->info) return osqp_error(OSQP_WORKSPACE_NOT_INIT_ERROR);
  if (!solution) return osqp_error(OSQP_WORKSPACE_NOT_INIT_ERROR);
  store_solution(solver, solution);
  return OSQP_NO_ERROR;
}

/* src/osqp_api.c:1057-1174 */
OSQPInt osqp_cleanup(OSQPSolver* solver) {
  OSQPWorkspace* work;
  if (!solver) return 0;
  work = solver->work;
  if (work) {
    if (work->data) {
      OSQPMatrix_free(work->data->P);
      OSQPMatrix_free(work->data->A);
      OSQPVectorf_free(work->data->q);
      OSQPVectorf_free(work->data->l);
      OSQPVectorf_free(work->data->u);
      free(work->data);
    }
    if (work->scaling) {
      OSQPVectorf_free(work->scaling->D);
      OSQPVectorf_free(work->scaling->Dinv);
      OSQPVectorf_free(work->scaling->E);
      OSQPVectorf_free(work->scaling->Einv);
    }
    free(work->scaling);
    OSQPVectorf_free(work->D_temp);
    OSQPVectorf_free(work->D_temp_A);
    OSQPVectorf_free(work->E_temp);
    if (work->linsys_solver && work->linsys_solver->free) work->linsys_solver->free(work->linsys_solver);
    OSQPVectorf_free(work->rho_vec);
    OSQPVectorf_free(work->rho_inv_vec);
    OSQPVectori_free(work->constr_type);
    OSQPVectorf_free(work->x);
    OSQPVectorf_free(work->z);
    OSQPVectorf_free(work->xz_tilde);
    OSQPVectorf_view_free(work->xtilde_view);
    OSQPVectorf_view_free(work->ztilde_view);
    OSQPVectorf_free(work->x_prev);
    OSQPVectorf_free(work->z_prev);
    OSQPVectorf_free(work->y);
    OSQPVectorf_free(work->Ax);
    OSQPVectorf_free(work->Px);
    OSQPVectorf_free(work->Aty);
    OSQPVectorf_free(work->delta_y);
    OSQPVectorf_free(work->Atdelta_y);
    OSQPVectorf_free(work->delta_x);
    OSQPVectori_free(work->pol_active);
    OSQPVectorf_free(work->pol_x);
    OSQPVectorf_free(work->pol_z);
    OSQPVectorf_free(work->pol_y);
    OSQPVectorf_free(work->Pdelta_x);
    OSQPVectorf_free(work->Adelta_x);
    if (solver->settings) free(solver->settings);
    if (solver->solution) {
      free(solver->solution->x);
      free(solver->solution->y);
      free(solver->solution->prim_inf_cert);
      free(solver->solution->dual_inf_cert);
      free(solver->solution);
    }
    if (solver->info) free(solver->info);
    free(work);
    /* the backend is reference-counted: this releases only this solver's share */
    osqp_algebra_free_libs();
  }
  free(solver);
  return 0;
}

/* src/osqp_api.c:1184-1257 */
OSQPInt osqp_update_data_vec(OSQPSolver* solver, const OSQPFloat* q_new, const OSQPFloat* l_new,
                             const OSQPFloat* u_new) {
  OSQPInt exitflag = 0;
  OSQPVectorf *l_tmp, *u_tmp;
  OSQPWorkspace* work;
  OSQPFloat t0;
  if (!solver || !solver->work) return osqp_error(OSQP_WORKSPACE_NOT_INIT_ERROR);
  work = solver->work;
  if (work->clear_update_time == 1) {
    work->clear_update_time = 0;
    solver->info->update_time = 0.0;
  }
  t0 = osqp_amd_toc(work);
  if (l_new || u_new) {
    l_tmp = work->z_prev;
    u_tmp = work->delta_y;
    if (l_new) OSQPVectorf_from_raw(l_tmp, l_new);
    if (u_new) OSQPVectorf_from_raw(u_tmp, u_new);
    if (solver->settings->scaling) {
      if (l_new) OSQPVectorf_ew_prod(l_tmp, l_tmp, work->scaling->E);
      if (u_new) OSQPVectorf_ew_prod(u_tmp, u_tmp, work->scaling->E);
    }
    if (l_new && u_new) exitflag = !OSQPVectorf_all_leq(l_tmp, u_tmp);
    else if (l_new) exitflag = !OSQPVectorf_all_leq(l_tmp, work->data->u);
    else exitflag = !OSQPVectorf_all_leq(work->data->l, u_tmp);
    if (exitflag) return osqp_error(OSQP_DATA_VALIDATION_ERROR);
    if (l_new) swap_vectors(&work->z_prev, &work->data->l);
    if (u_new) swap_vectors(&work->delta_y, &work->data->u);
    if (solver->settings->rho_is_vec) exitflag = update_rho_vec(solver);
  }
  if (q_new) {
    OSQPVectorf_from_raw(work->data->q, q_new);
    if (solver->settings->scaling) {
      OSQPVectorf_ew_prod(work->data->q, work->data->q, work->scaling->D);
      OSQPVectorf_mult_scalar(work->data->q, work->scaling->c);
    }
  }
  reset_info(solver->info);
  solver->info->update_time += osqp_amd_toc(work) - t0;
  return exitflag;
}

/* src/osqp_api.c:1260-1293 */
OSQPInt osqp_warm_start(OSQPSolver* solver, const OSQPFloat* x, const OSQPFloat* y) {
  OSQPWorkspace* work;
  if (!solver || !solver->work) return osqp_error(OSQP_WORKSPACE_NOT_INIT_ERROR);
  work = solver->work;
  if (!solver->settings->warm_starting) solver->settings->warm_starting = 1;
  if (x) OSQPVectorf_from_raw(work->x, x);
  if (y) OSQPVectorf_from_raw(work->y, y);
  if (solver->settings->scaling) {
    if (x) OSQPVectorf_ew_prod(work->x, work->x, work->scaling->Dinv);
    if (y) {
      OSQPVectorf_ew_prod(work->y, work->y, work->scaling->Einv);
      OSQPVectorf_mult_scalar(work->y, work->scaling->c);
    }
  }
  if (x) OSQPMatrix_Axpy(work->data->A, work->x, work->z, 1.0, 0.0);
  work->linsys_solver->warm_start(work->linsys_solver, work->x);
  return 0;
}

/* src/osqp_api.c:1296-1304 */
void osqp_cold_start(OSQPSolver* solver) {
  OSQPWorkspace* work = solver->work;
  OSQPVectorf_set_scalar(work->x, 0.);
  OSQPVectorf_set_scalar(work->z, 0.);
  OSQPVectorf_set_scalar(work->y, 0.);
  work->linsys_solver->warm_start(work->linsys_solver, work->x);
}

/* src/osqp_api.c:1309-1409 */
OSQPInt osqp_update_data_mat(OSQPSolver* solver, const OSQPFloat* Px_new, const OSQPInt* Px_new_idx,
                             OSQPInt P_new_n, const OSQPFloat* Ax_new, const OSQPInt* Ax_new_idx, OSQPInt A_new_n) {
  OSQPInt exitflag, nnzP, nnzA;
  OSQPWorkspace* work;
  OSQPFloat t0;
  if (!solver || !solver->work) return osqp_error(OSQP_WORKSPACE_NOT_INIT_ERROR);
  work = solver->work;
  if (work->clear_update_time == 1) {
    work->clear_update_time = 0;
    solver->info->update_time = 0.0;
  }
  t0 = osqp_amd_toc(work);
  nnzP = OSQPMatrix_get_nz(work->data->P);
  nnzA = OSQPMatrix_get_nz(work->data->A);
  if (P_new_n > nnzP || P_new_n < 0) {
    fprintf(stderr, "new number of elements (%i) out of bounds for P (%i max)\n", (int)P_new_n, (int)nnzP);
    return 1;
  }
  if (Px_new_idx == NULL && P_new_n != 0 && P_new_n != nnzP) {
    fprintf(stderr, "index vector is required for partial updates of P\n");
    return 1;
  }
  if (P_new_n == 0) P_new_n = nnzP;
  if (A_new_n > nnzA || A_new_n < 0) {
    fprintf(stderr, "new number of elements (%i) out of bounds for A (%i max)\n", (int)A_new_n, (int)nnzA);
    return 2;
  }
  if (Ax_new_idx == NULL && A_new_n != 0 && A_new_n != nnzA) {
    fprintf(stderr, "index vector is required for partial updates of A\n");
    return 2;
  }
  if (A_new_n == 0) A_new_n = nnzA;
  if (solver->settings->scaling) unscale_data(solver);
  if (Px_new) OSQPMatrix_update_values(work->data->P, Px_new, Px_new_idx, P_new_n);
  if (Ax_new) OSQPMatrix_update_values(work->data->A, Ax_new, Ax_new_idx, A_new_n);
  if (solver->settings->scaling) scale_data(solver);
  if (solver->settings->scaling)
    exitflag = work->linsys_solver->update_matrices(work->linsys_solver, work->data->P, NULL, nnzP, work->data->A,
                                                    NULL, nnzA);
  else
    exitflag = work->linsys_solver->update_matrices(work->linsys_solver, work->data->P, Px_new_idx, P_new_n,
                                                    work->data->A, Ax_new_idx, A_new_n);
  reset_info(solver->info);
  if (exitflag != 0) fprintf(stderr, "new KKT matrix is not quasidefinite\n");
  solver->info->update_time += osqp_amd_toc(work) - t0;
  return exitflag;
}

/* src/osqp_api.c:1412-1464 */
OSQPInt osqp_update_rho(OSQPSolver* solver, OSQPFloat rho_new) {
  OSQPInt exitflag;
  OSQPWorkspace* work;
  OSQPFloat t0 = 0.0;
  if (!solver || !solver->work) return osqp_error(OSQP_WORKSPACE_NOT_INIT_ERROR);
  work = solver->work;
  if (rho_new <= 0) {
    fprintf(stderr, "rho must be positive\n");
    return 1;
  }
  if (work->rho_update_from_solve == 0) {
    if (work->clear_update_time == 1) {
      work->clear_update_time = 0;
      solver->info->update_time = 0.0;
    }
    t0 = osqp_amd_toc(work);
  }
  solver->settings->rho = c_min(c_max(rho_new, OSQP_RHO_MIN), OSQP_RHO_MAX);
  if (solver->settings->rho_is_vec) {
    OSQPVectorf_set_scalar_conditional(work->rho_vec, work->constr_type, OSQP_RHO_MIN, solver->settings->rho,
                                       OSQP_RHO_EQ_OVER_RHO_INEQ * solver->settings->rho);
    OSQPVectorf_ew_reciprocal(work->rho_inv_vec, work->rho_vec);
  } else {
    work->rho_inv = 1. / solver->settings->rho;
  }
  exitflag = work->linsys_solver->update_rho_vec(work->linsys_solver, work->rho_vec, solver->settings->rho);
  if (work->rho_update_from_solve == 0) solver->info->update_time += osqp_amd_toc(work) - t0;
  return exitflag;
}

/* src/osqp_api.c:1474-1530 */
OSQPInt osqp_update_settings(OSQPSolver* solver, const OSQPSettings* new_settings) {
  OSQPSettings* settings;
  if (!solver || !solver->work) return osqp_error(OSQP_WORKSPACE_NOT_INIT_ERROR);
  settings = solver->settings;
  if (validate_settings(new_settings, 0)) return osqp_error(OSQP_SETTINGS_VALIDATION_ERROR);
  settings->profiler_level = new_settings->profiler_level;
  osqp_amd_profiler_level(settings->profiler_level);  /* src/osqp_api.c:1490-1491 */
  settings->verbose = new_settings->verbose;
  settings->warm_starting = new_settings->warm_starting;
  settings->polishing = new_settings->polishing;
  settings->alpha = new_settings->alpha;
  settings->cg_max_iter = new_settings->cg_max_iter;
  settings->cg_tol_reduction = new_settings->cg_tol_reduction;
  settings->cg_tol_fraction = new_settings->cg_tol_fraction;
  settings->cg_precond = new_settings->cg_precond;
  settings->max_iter = new_settings->max_iter;
  settings->eps_abs = new_settings->eps_abs;
  settings->eps_rel = new_settings->eps_rel;
  settings->eps_prim_inf = new_settings->eps_prim_inf;
  settings->eps_dual_inf = new_settings->eps_dual_inf;
  settings->scaled_termination = new_settings->scaled_termination;
  settings->check_termination = new_settings->check_termination;
  settings->check_dualgap = new_settings->check_dualgap;
  settings->time_limit = new_settings->time_limit;
  settings->delta = new_settings->delta;
  settings->polish_refine_iter = new_settings->polish_refine_iter;
  solver->work->linsys_solver->update_settings(solver->work->linsys_solver, settings);
  return 0;
}

OSQPInt osqp_amd_solver_stats(OSQPSolver* solver, OSQPInt* pcg_iters_total, OSQPInt* linsys_solves, double* gpu_ms) {
  if (!solver || !solver->work || !solver->work->linsys_solver) return 1;
  osqp_amd_linsys_counters(solver->work->linsys_solver, pcg_iters_total, linsys_solves);
  if (gpu_ms) *gpu_ms = solver->info->solve_time * 1e3;
  return 0;
}

OSQPInt osqp_amd_time_pcg_kernels(OSQPSolver* solver, OSQPInt reps, double* ms_av, double* ms_kv, double* bytes_av,
                                  double* bytes_kv) {
  if (!solver || !solver->work || !solver->work->linsys_solver) return 1;
  return osqp_amd_linsys_time_kernels(solver->work->linsys_solver, reps, ms_av, ms_kv, bytes_av, bytes_kv);
}

OSQPInt osqp_amd_apply_operator(OSQPSolver* solver, OSQPInt which, const OSQPFloat* x, OSQPFloat* y) {
  if (!solver || !solver->work || !solver->work->linsys_solver) return 1;
  return osqp_amd_linsys_apply(solver->work->linsys_solver, which, x, y);
}

OSQPInt osqp_amd_time_fp64_kernels(OSQPSolver* solver, OSQPInt reps, double* ms_a, double* ms_k, double* bytes_a,
                                   double* bytes_k) {
  if (!solver || !solver->work || !solver->work->linsys_solver) return 1;
  return osqp_amd_linsys_time_fp64_kernels(solver->work->linsys_solver, reps, ms_a, ms_k, bytes_a, bytes_k);
}

OSQPInt osqp_amd_pcg_bytes(OSQPSolver* solver, double* per_cg_iter, double* per_solve) {
  if (!solver || !solver->work || !solver->work->linsys_solver) return 1;
  return osqp_amd_linsys_bytes(solver->work->linsys_solver, per_cg_iter, per_solve);
}

/* The vector passes of update_info (auxil.c), in 8-byte element reads and
 * writes per element of n and of m: residuals, objective terms and the
 * tolerance norms (the scaled norms of unscaled termination read two
 * vectors); the products A x, P x and A' y are counted separately. */
OSQPInt osqp_amd_check_bytes(OSQPSolver* solver, double* bytes) {
  if (!solver || !solver->work || !solver->work->data || !bytes) return 1;
  const OSQPWorkspace* work = solver->work;
  const double n = work->data->n, m = work->data->m;
  const int uns = solver->settings->scaling && !solver->settings->scaled_termination;
  double b = osqp_amd_matrix_product_bytes(work->data->P, 0);
  if (m > 0) b += osqp_amd_matrix_product_bytes(work->data->A, 0) + osqp_amd_matrix_product_bytes(work->data->A, 1);
  b += 8.0 * (uns ? 21.0 * n + 22.0 * m : 16.0 * n + 18.0 * m);
  *bytes = b;
  return 0;
}

OSQPInt osqp_amd_solver_matrices(OSQPSolver* solver, const OSQPMatrix** P, const OSQPMatrix** A) {
  if (!solver || !solver->work || !solver->work->data) return 1;
  if (P) *P = solver->work->data->P;
  if (A) *A = solver->work->data->A;
  return 0;
}

OSQPInt osqp_amd_solver_scaling(OSQPSolver* solver, OSQPFloat* D, OSQPFloat* E, OSQPFloat* c, OSQPFloat* q,
                                OSQPFloat* l, OSQPFloat* u) {
  OSQPWorkspace* work;
  if (!solver || !solver->work || !solver->work->data) return 1;
  work = solver->work;
  if (q) OSQPVectorf_to_raw(q, work->data->q);
  if (l && work->data->m) OSQPVectorf_to_raw(l, work->data->l);
  if (u && work->data->m) OSQPVectorf_to_raw(u, work->data->u);
  if (!work->scaling) return 1;
  if (D) OSQPVectorf_to_raw(D, work->scaling->D);
  if (E && work->data->m) OSQPVectorf_to_raw(E, work->scaling->E);
  if (c) *c = work->scaling->c;
  return 0;
}
