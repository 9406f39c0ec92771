/*
 * osqp_amd_algebra.h — the algebra/hip backend: OSQP's algebra plugin
 * interface implemented with hand-written gfx950 HIP kernels.
 *
 * These are the symbols OSQP's solver core (src/osqp_api.c, src/auxil.c,
 * src/scaling.c) calls; a maintainer adds this backend to the reference build
 * as algebra/hip and links libosqp_amd.so instead of algebra/builtin.
 * Declarations and their meaning follow:
 *   include/private/algebra_vector.h:28-290
 *   include/private/algebra_matrix.h:26-138
 *   include/private/lin_alg.h:13-53
 *   include/private/types.h:243-279   (struct linsys_solver)
 *
 * All vector and matrix storage lives in device memory (HBM); vectors are
 * FP64, indices int32.  Functions returning a scalar synchronise the backend
 * stream; everything else is asynchronous on it.
 */
#ifndef OSQP_AMD_ALGEBRA_H
#define OSQP_AMD_ALGEBRA_H

#include <stddef.h>

#include "osqp_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OSQPVectorf_ OSQPVectorf;
typedef struct OSQPVectori_ OSQPVectori;
typedef struct OSQPMatrix_  OSQPMatrix;
typedef struct linsys_solver LinSysSolver;

/* ---- vectors: include/private/algebra_vector.h ------------------------ */
OSQPInt      OSQPVectorf_is_eq(const OSQPVectorf* A, const OSQPVectorf* B, OSQPFloat tol); /* :28 */
OSQPVectorf* OSQPVectorf_new(const OSQPFloat* a, OSQPInt length);                          /* :39 */
OSQPVectori* OSQPVectori_new(const OSQPInt* a, OSQPInt length);                            /* :43 */
OSQPVectorf* OSQPVectorf_malloc(OSQPInt length);                                           /* :33 */
OSQPVectori* OSQPVectori_malloc(OSQPInt length);                                           /* :35 */
OSQPVectorf* OSQPVectorf_calloc(OSQPInt length);                                           /* :34 */
OSQPVectori* OSQPVectori_calloc(OSQPInt length);                                           /* :36 */
OSQPVectorf* OSQPVectorf_copy_new(const OSQPVectorf* a);                                   /* :47 */
void         OSQPVectorf_free(OSQPVectorf* a);                                             /* :50 */
void         OSQPVectori_free(OSQPVectori* a);                                             /* :53 */
OSQPVectorf* OSQPVectorf_view(const OSQPVectorf* a, OSQPInt head, OSQPInt length);         /* :81 */
void         OSQPVectorf_view_free(OSQPVectorf* a);                                        /* :91 */
OSQPInt      OSQPVectorf_length(const OSQPVectorf* a);                                     /* :97 */
OSQPInt      OSQPVectori_length(const OSQPVectori* a);                                     /* :100 */
void OSQPVectorf_copy(OSQPVectorf* b, const OSQPVectorf* a);                               /* :106 */
void OSQPVectorf_from_raw(OSQPVectorf* b, const OSQPFloat* av);                            /* :110 */
void OSQPVectori_from_raw(OSQPVectori* b, const OSQPInt* av);                              /* :114 */
void OSQPVectorf_to_raw(OSQPFloat* bv, const OSQPVectorf* a);                              /* :118 */
void OSQPVectori_to_raw(OSQPInt* bv, const OSQPVectori* a);                                /* :122 */
void OSQPVectorf_set_scalar(OSQPVectorf* a, OSQPFloat sc);                                 /* :126 */
void OSQPVectorf_set_scalar_conditional(OSQPVectorf* a, const OSQPVectori* test,
                                        OSQPFloat sc_if_neg, OSQPFloat sc_if_zero,
                                        OSQPFloat sc_if_pos);                              /* :130 */
void OSQPVectorf_round_to_zero(OSQPVectorf* a, OSQPFloat tol);                             /* :139 */
void OSQPVectorf_mult_scalar(OSQPVectorf* a, OSQPFloat sc);                                /* :143 */
void OSQPVectorf_plus(OSQPVectorf* x, const OSQPVectorf* a, const OSQPVectorf* b);         /* :147 */
void OSQPVectorf_minus(OSQPVectorf* x, const OSQPVectorf* a, const OSQPVectorf* b);        /* :152 */
void OSQPVectorf_add_scaled(OSQPVectorf* x, OSQPFloat sca, const OSQPVectorf* a,
                            OSQPFloat scb, const OSQPVectorf* b);                          /* :157 */
void OSQPVectorf_add_scaled3(OSQPVectorf* x, OSQPFloat sca, const OSQPVectorf* a,
                             OSQPFloat scb, const OSQPVectorf* b,
                             OSQPFloat scc, const OSQPVectorf* c);                         /* :164 */
OSQPFloat OSQPVectorf_norm_inf(const OSQPVectorf* v);                                      /* :173 */
OSQPFloat OSQPVectorf_scaled_norm_inf(const OSQPVectorf* S, const OSQPVectorf* v);         /* :176 */
OSQPFloat OSQPVectorf_norm_inf_diff(const OSQPVectorf* a, const OSQPVectorf* b);           /* :180 */
OSQPFloat OSQPVectorf_norm_1(const OSQPVectorf* a);                                        /* :187 */
OSQPFloat OSQPVectorf_dot_prod(const OSQPVectorf* a, const OSQPVectorf* b);                /* :190 */
OSQPFloat OSQPVectorf_dot_prod_signed(const OSQPVectorf* a, const OSQPVectorf* b,
                                      OSQPInt sign);                                       /* :198 */
void OSQPVectorf_ew_prod(OSQPVectorf* c, const OSQPVectorf* a, const OSQPVectorf* b);      /* :203 */
OSQPInt OSQPVectorf_all_leq(const OSQPVectorf* l, const OSQPVectorf* u);                   /* :210 */
void OSQPVectorf_ew_bound_vec(OSQPVectorf* x, const OSQPVectorf* z,
                              const OSQPVectorf* l, const OSQPVectorf* u);                 /* :217 */
void OSQPVectorf_project_polar_reccone(OSQPVectorf* y, const OSQPVectorf* l,
                                       const OSQPVectorf* u, OSQPFloat infval);            /* :227 */
OSQPInt OSQPVectorf_in_reccone(const OSQPVectorf* y, const OSQPVectorf* l,
                               const OSQPVectorf* u, OSQPFloat infval, OSQPFloat tol);     /* :237 */
void OSQPVectorf_ew_reciprocal(OSQPVectorf* b, const OSQPVectorf* a);                      /* :246 */
void OSQPVectorf_ew_sqrt(OSQPVectorf* a);                                                  /* :250 */
void OSQPVectorf_ew_max_vec(OSQPVectorf* c, const OSQPVectorf* a, const OSQPVectorf* b);   /* :253 */
void OSQPVectorf_ew_min_vec(OSQPVectorf* c, const OSQPVectorf* a, const OSQPVectorf* b);   /* :258 */
OSQPInt OSQPVectorf_ew_bounds_type(OSQPVectori* iseq, const OSQPVectorf* l,
                                   const OSQPVectorf* u, OSQPFloat tol, OSQPFloat infval); /* :269 */
void OSQPVectorf_set_scalar_if_lt(OSQPVectorf* x, const OSQPVectorf* z,
                                  OSQPFloat testval, OSQPFloat newval);                    /* :279 */
void OSQPVectorf_set_scalar_if_gt(OSQPVectorf* x, const OSQPVectorf* z,
                                  OSQPFloat testval, OSQPFloat newval);                    /* :287 */

/* ---- matrices: include/private/algebra_matrix.h ----------------------- */
OSQPInt     OSQPMatrix_is_eq(const OSQPMatrix* A, const OSQPMatrix* B, OSQPFloat tol);     /* :26 */
OSQPMatrix* OSQPMatrix_new_from_csc(const OSQPCscMatrix* A, OSQPInt is_triu);              /* :34 */
void        OSQPMatrix_update_values(OSQPMatrix* M, const OSQPFloat* Mx_new,
                                     const OSQPInt* Mx_new_idx, OSQPInt M_new_n);          /* :60 */
OSQPInt     OSQPMatrix_get_m(const OSQPMatrix* M);                                         /* :66 */
OSQPInt     OSQPMatrix_get_n(const OSQPMatrix* M);                                         /* :69 */
OSQPInt     OSQPMatrix_get_nz(const OSQPMatrix* M);                                        /* :81 */
void        OSQPMatrix_mult_scalar(OSQPMatrix* A, OSQPFloat sc);                           /* :86 */
void        OSQPMatrix_lmult_diag(OSQPMatrix* A, const OSQPVectorf* L);                    /* :90 */
void        OSQPMatrix_rmult_diag(OSQPMatrix* A, const OSQPVectorf* R);                    /* :94 */
void        OSQPMatrix_Axpy(const OSQPMatrix* A, const OSQPVectorf* x, OSQPVectorf* y,
                            OSQPFloat alpha, OSQPFloat beta);                              /* :107 */
void        OSQPMatrix_Atxpy(const OSQPMatrix* A, const OSQPVectorf* x, OSQPVectorf* y,
                             OSQPFloat alpha, OSQPFloat beta);                             /* :114 */
void        OSQPMatrix_col_norm_inf(const OSQPMatrix* M, OSQPVectorf* E);                  /* :125 */
void        OSQPMatrix_row_norm_inf(const OSQPMatrix* M, OSQPVectorf* E);                  /* :128 */
void        OSQPMatrix_free(OSQPMatrix* M);                                                /* :135 */
OSQPMatrix* OSQPMatrix_submatrix_byrows(const OSQPMatrix* A, const OSQPVectori* rows);     /* :137 */
/* d = diag(A' diag(D) A)  (algebra_matrix.h:98) */
void        OSQPMatrix_AtDA_extract_diag(const OSQPMatrix* A, const OSQPVectorf* D,
                                         OSQPVectorf* d);
/* d = diag(A)  (algebra_matrix.h:103) */
void        OSQPMatrix_extract_diag(const OSQPMatrix* A, OSQPVectorf* d);

/* ---- libraries and linear system: include/private/lin_alg.h ---------- */
OSQPInt osqp_algebra_linsys_supported(void);                                               /* :13 */
enum osqp_linsys_solver_type osqp_algebra_default_linsys(void);                            /* :16 */
OSQPInt osqp_algebra_init_libs(OSQPInt device);                                            /* :19 */
void    osqp_algebra_free_libs(void);                                                      /* :22 */
OSQPInt osqp_algebra_name(char* name, OSQPInt nameLen);                                    /* :25 */
OSQPInt osqp_algebra_device_name(char* name, OSQPInt nameLen);                             /* :28 */
OSQPInt osqp_algebra_init_linsys_solver(LinSysSolver** s, const OSQPMatrix* P,
                                        const OSQPMatrix* A, const OSQPVectorf* rho_vec,
                                        const OSQPSettings* settings,
                                        OSQPFloat* scaled_prim_res,
                                        OSQPFloat* scaled_dual_res,
                                        OSQPInt polishing);                                /* :46 */

/* include/private/types.h:243-279 — the linear-system plugin vtable */
struct linsys_solver {
  enum osqp_linsys_solver_type type;
  const char* (*name)(LinSysSolver* self);
  OSQPInt (*solve)(LinSysSolver* self, OSQPVectorf* b, OSQPInt admm_iter);
  void    (*update_settings)(LinSysSolver* self, const OSQPSettings* settings);
  void    (*warm_start)(LinSysSolver* self, const OSQPVectorf* x);
  OSQPInt (*adjoint_derivative)(LinSysSolver* self);
  void    (*free)(LinSysSolver* self);
  OSQPInt (*update_matrices)(LinSysSolver* self,
                             const OSQPMatrix* P, const OSQPInt* Px_new_idx, OSQPInt P_new_n,
                             const OSQPMatrix* A, const OSQPInt* Ax_new_idx, OSQPInt A_new_n);
  OSQPInt (*update_rho_vec)(LinSysSolver* self, const OSQPVectorf* rho_vec, OSQPFloat rho_sc);
  OSQPInt nthreads;
};

/* ---- MI355X extensions (no counterpart in the reference) -------------- */
/* Device pointer of a vector's storage (for zero-copy hand-off between
 * the host core and the fused kernels of this backend). */
OSQPFloat* OSQPVectorf_device_ptr(const OSQPVectorf* a);
/* Blocks until all queued work of the backend stream has finished. */
void       osqp_amd_algebra_sync(void);
/* Last-solve counters of a PCG linsys: total CG iterations and solve calls. */
void       osqp_amd_linsys_counters(LinSysSolver* s, OSQPInt* cg_iters, OSQPInt* solves);
/* Mean duration (ms) of one launch of each CG operator kernel on this
 * linsys' data (w = rho o A p; K p = sigma p + P p + A' w) over `reps`
 * launches timed with HIP events, and the algorithmic bytes of one launch. */
OSQPInt    osqp_amd_linsys_time_kernels(LinSysSolver* s, OSQPInt reps, double* ms_av, double* ms_kv,
                                        double* bytes_av, double* bytes_kv);
/* The same for the FP64 panel SpMVs of the defect correction (z = A x and
 * the residual's [P A'] [x; w]); returns 1 when those operators are not
 * panelled.  Extension: the reference has one precision per build. */
OSQPInt    osqp_amd_linsys_time_fp64_kernels(LinSysSolver* s, OSQPInt reps, double* ms_a, double* ms_k,
                                             double* bytes_a, double* bytes_k);
/* Applies one of the linsys' operators to the host vector x through the
 * kernels its solves use (column panels, hybrids or row-block CSR), for checks
 * of each operator form against a reference product (no counterpart in the
 * reference):
 *   which 0: y (n) = (P + sigma I + A' diag(rho) A) x in the CG precision;
 *   which 1: y (n) = the same with the defect correction's FP64-vector operator;
 *   which 2: y (m) = A x, the defect correction's z = A x product.
 * P, A and rho are the linsys' own (scaled) data.  Resets the CG state.
 * Returns 0, or 1 for a bad argument or a HIP error. */
OSQPInt    osqp_amd_linsys_apply(LinSysSolver* s, OSQPInt which, const OSQPFloat* x, OSQPFloat* y);
/* Algorithmic bytes (each array element touched once per pass) of one CG
 * iteration and of the per-solve work outside the CG loop (cached A x path:
 * w = rho o (A x - b2), FP64 residual, correction, z = A x) of this linsys.
 * With the incremental z of the FP32 engine, both are averages over its
 * refresh cycle (one exact z = A x64 product in eight solves). */
OSQPInt    osqp_amd_linsys_bytes(LinSysSolver* s, double* per_cg_iter, double* per_solve);
/* Algorithmic bytes of one product y = M x (transpose 0) or y = M' x
 * (transpose 1) on the kernel path the matrix takes now (FP64 panels or
 * row-block CSR), y written, not read.  No counterpart in the reference: a
 * measurement helper. */
double     osqp_amd_matrix_product_bytes(const OSQPMatrix* M, OSQPInt transpose);
/* Fused ADMM steps for the rho-vector case (no counterpart in the reference,
 * which issues the same arithmetic as separate vector calls; results are
 * bit-identical): compute_rhs (src/auxil.c:136-158) and update_x / update_z /
 * update_y (src/auxil.c:172-229) in one pass each.  With next = 1 the update
 * also writes the next iteration's compute_rhs result (from the new x, z, y,
 * with sigma and q) over xt / zt, valid while rho, sigma and q stay; w_next
 * (device pointer, may be null) receives rho o (zt_read - zt_next). */
void       osqp_amd_admm_rhs(OSQPVectorf* xt, OSQPVectorf* zt, OSQPFloat sigma, const OSQPVectorf* x_prev,
                             const OSQPVectorf* q, const OSQPVectorf* rho_inv, const OSQPVectorf* y,
                             const OSQPVectorf* z_prev);
void       osqp_amd_admm_update(OSQPFloat alpha, OSQPVectorf* x, OSQPVectorf* delta_x, OSQPVectorf* xt,
                                const OSQPVectorf* x_prev, OSQPVectorf* z, OSQPVectorf* y, OSQPVectorf* delta_y,
                                OSQPVectorf* zt, const OSQPVectorf* z_prev, const OSQPVectorf* rho,
                                const OSQPVectorf* rho_inv, const OSQPVectorf* l, const OSQPVectorf* u,
                                OSQPInt next, OSQPFloat sigma, const OSQPVectorf* q, OSQPFloat* w_next);
/* The PCG's hand-off for that fused right-hand side (no counterpart in the
 * reference): w_target returns the device m-vector in which the update may
 * write w = rho o (A x - b2) for the next solve (A x being the z~ the last
 * solve returned), or null when the next solve would not use it;
 * set_w_given tells the next solve whether it was written. */
OSQPFloat* osqp_amd_linsys_w_target(LinSysSolver* s);
void       osqp_amd_linsys_set_w_given(LinSysSolver* s, OSQPInt given);
/* Queued reductions (no counterpart in the reference, whose reductions each
 * return to the host): each call queues the same reduction as
 * OSQPVectorf_norm_inf / scaled_norm_inf / dot_prod / dot_prod_signed (same
 * arithmetic, same bits) and returns at once; osqp_amd_red_wait() runs all
 * queued ones in one launch with one host round trip and stores each result
 * in its *dst.  A queued reduction reads its vectors when the queue runs, so
 * queue it only once its inputs are final.  Up to 16 may be queued (a
 * further call runs the queue first). */
void       osqp_amd_red_norm_inf(const OSQPVectorf* v, OSQPFloat* dst);
void       osqp_amd_red_scaled_norm_inf(const OSQPVectorf* S, const OSQPVectorf* v, OSQPFloat* dst);
void       osqp_amd_red_dot_prod(const OSQPVectorf* a, const OSQPVectorf* b, OSQPFloat* dst);
void       osqp_amd_red_dot_prod_signed(const OSQPVectorf* a, const OSQPVectorf* b, OSQPInt sign, OSQPFloat* dst);
void       osqp_amd_red_wait(void);
/* Reductions that allocate nothing.  No counterpart in the reference: this is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It runs
 * 1..16 reductions, each as the queued osqp_amd_red_* call of the same kind
 * runs it (same kernel, same bits), in one launch on the caller's `stream`
 * (a hipStream_t; null is the default stream), and writes reduction k's
 * value to out[k], out being device memory of `count` doubles.  a and b are
 * device pointers of n doubles; n = 0 gives 0.  The per-block partials and
 * tickets live in a fixed 64.1 KiB device array per GPU, held until the
 * process ends: calls that overlap on one GPU race.  Returns 0, or
 * OSQP_DATA_VALIDATION_ERROR (count outside 1..16, unknown kind, negative n,
 * null operand) without launching anything.  Needs no osqp_algebra_init_libs. */
enum osqp_amd_reduction_kind {
  OSQP_AMD_RED_NORM_INF = 0,     /* max |a_i|            (OSQPVectorf_norm_inf)          */
  OSQP_AMD_RED_SCALED_NORM_INF,  /* max |a_i b_i|        (OSQPVectorf_scaled_norm_inf)   */
  OSQP_AMD_RED_NORM_INF_DIFF,    /* max |a_i - b_i|      (OSQPVectorf_norm_inf_diff)     */
  OSQP_AMD_RED_NORM_1,           /* sum |a_i|            (OSQPVectorf_norm_1)            */
  OSQP_AMD_RED_DOT,              /* sum a_i b_i          (OSQPVectorf_dot_prod)          */
  OSQP_AMD_RED_DOT_POS,          /* sum a_i max(b_i, 0)  (OSQPVectorf_dot_prod_signed 1) */
  OSQP_AMD_RED_DOT_NEG           /* sum a_i min(b_i, 0)  (OSQPVectorf_dot_prod_signed -1)*/
};
typedef struct {
  OSQPInt kind;  /* enum osqp_amd_reduction_kind */
  OSQPInt n;
  const OSQPFloat* a;
  const OSQPFloat* b;  /* unused by NORM_INF and NORM_1 */
} OSQPAmdReduction;
OSQPInt    osqp_amd_reduce_noalloc(const OSQPAmdReduction* items, OSQPInt count, OSQPFloat* out, void* stream);
/* Last HIP error seen by the backend (0 if none); cleared on read. */
OSQPInt    osqp_amd_last_error(void);
/* 1 when p points to device memory (hipMalloc), 0 for host memory, and a
 * copy of bytes from either to host memory (0 on success).  osqp_setup and
 * the serving engine accept device-resident arrays through them (the
 * reference's setup reads host arrays only, src/auxil.c:950-1017). */
OSQPInt    osqp_amd_ptr_is_device(const void* p);
OSQPInt    osqp_amd_copy_to_host(void* dst, const void* src, size_t bytes);
/* Copies the backend's structures of M to host arrays, for checks of the
 * integer work against OSQP's (csc_utils.c:247-286 triplet_to_csr,
 * cuda_csr.cu:489-577 triu->full and transpose).  Any pointer may be null.
 *   rp, ci, val : the CSR copy (A: m+1 / nnz entries; P: the full symmetric
 *                 matrix, n+1 / nnz_csr entries, rows sorted by column);
 *   trp, tci, tval: CSR of A' (n+1 / nnz entries; general matrices only);
 *   csc2csr     : CSR slot of each entry of the input CSC (nnz_in);
 *   csc2mirror  : P only, the transposed slot of each input entry, -1 on
 *                 the diagonal (nnz_in);
 *   diag        : P only, the slot of (j, j) in row j, -1 if absent (n).
 * *nnz_csr / *nnz_in (may be null) receive the two counts.  Returns 0. */
OSQPInt    osqp_amd_matrix_export(const OSQPMatrix* M, OSQPInt* nnz_csr, OSQPInt* nnz_in, OSQPInt* rp, OSQPInt* ci,
                                  OSQPFloat* val, OSQPInt* trp, OSQPInt* tci, OSQPFloat* tval, OSQPInt* csc2csr,
                                  OSQPInt* csc2mirror, OSQPInt* diag);
/* The solver's problem matrices (scaled, as the ADMM iterations see them). */
OSQPInt    osqp_amd_solver_matrices(OSQPSolver* solver, const OSQPMatrix** P, const OSQPMatrix** A);
/* The solver's Ruiz scaling (src/scaling.c:48-164) and scaled vectors:
 * D (n), E (m), c, and the scaled q (n), l, u (m).  Any pointer may be null;
 * returns 1 when the solver was set up without scaling. */
OSQPInt    osqp_amd_solver_scaling(OSQPSolver* solver, OSQPFloat* D, OSQPFloat* E, OSQPFloat* c, OSQPFloat* q,
                                   OSQPFloat* l, OSQPFloat* u);

#ifdef __cplusplus
}
#endif

#endif /* OSQP_AMD_ALGEBRA_H */
