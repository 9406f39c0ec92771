/*
 * osqp_amd_serve.h — batched serving engine for many small QPs on one GPU.
 *
 * No counterpart in the reference: OSQP solves one problem per solver object
 * (include/public/osqp_api_functions.h osqp_setup / osqp_solve).  This engine
 * keeps many problems resident in HBM and advances all of them together:
 *
 *   - a problem is added with the same data osqp_setup takes (CSC P upper
 *     triangle, q, CSC A, l, u) and is validated the same way; it is admitted
 *     as soon as a slot and device memory are free, then scaled, set up and
 *     solved on the GPU by one workgroup, with the host core's ADMM semantics
 *     and the PCG linear solver (FP32 CG with FP64 defect correction, or FP64);
 *   - problems run in epochs: in one epoch every active problem of a size
 *     bucket advances up to a fixed number of ADMM iterations in one
 *     persistent kernel (one hipGraph per bucket, each bucket on its own
 *     stream); a workgroup that finishes a problem takes the next one in the
 *     same launch; problems added meanwhile join their bucket's next epoch.
 *     Steps are pipelined: a step finishes the buckets whose epochs have
 *     ended (waiting for at least one when any is in flight), launches the
 *     next epoch of every idle bucket with work and returns, so the caller's
 *     work between steps overlaps the GPU and a solve is reported by the step
 *     that finishes the epoch it ended in;
 *   - a resident problem keeps its iterates and rho after a solve; a new q
 *     (osqp_update_data_vec semantics) and a warm-started re-solve are
 *     requested with osqp_batch_resolve.
 *
 * A problem's results do not depend on the other problems in flight or on the
 * engine's tuning (team sizes): its arithmetic is fixed by its own data.
 *
 * Settings: those of OSQPSettings, with rho_is_vec = 1 and polishing = 0.
 * All four adaptive_rho modes apply (src/osqp_api.c:776-835); the time-based
 * one measures the setup and loop times on the device clock, so, as in the
 * reference, its results depend on timing.  time_limit (src/osqp_api.c:
 * 743-759) counts from the start of a problem's setup on the GPU (first
 * solve) or of its re-solve, on the device clock; limits of 1e6 s or more
 * are not applied (the default is 1e10 s).
 * Results follow OSQP's contract (src/auxil.c:598-674): x and y are OSQP_NAN
 * without a solution, and the infeasibility certificates are returned
 * normalised (OSQP_NAN where they do not apply).
 * Calls on one engine must not overlap (one host thread drives an engine).
 */
#ifndef OSQP_AMD_SERVE_H
#define OSQP_AMD_SERVE_H

#include "osqp_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OSQPBatch OSQPBatch;

typedef struct {
  OSQPInt   status_val;
  OSQPInt   iter;
  OSQPInt   rho_updates;
  OSQPFloat obj_val;
  OSQPFloat dual_obj_val;
  OSQPFloat prim_res;
  OSQPFloat dual_res;
  OSQPFloat duality_gap;
  OSQPFloat rho_estimate;
  OSQPFloat primdual_int;
  OSQPFloat rel_kkt_error;
  OSQPFloat latency;      /* seconds from add / resolve to the result seen by the host */
  long long cg_iters;     /* CG iterations of this problem since it was added */
} OSQPBatchInfo;

/* precision: 32 (FP32 CG, FP64 residuals and ADMM state) or 64.
 * max_active: problems on the GPU at once (0: 4096).
 * arena_bytes: device memory for resident problems (0: 4 GiB).
 * Pinned host memory: a staging pool of arena_bytes / 4 (at least 64 MiB, at
 * most 1 GiB) for stage images, results and matrix updates, and an image arena
 * of half the pool's size where osqp_batch_add builds stage images (when it
 * cannot be had, images are built on the heap and copied at admission).
 * Returns 0, OSQP_SETTINGS_VALIDATION_ERROR or OSQP_MEM_ALLOC_ERROR. */
OSQPInt osqp_batch_create(OSQPBatch** batch, const OSQPSettings* settings, OSQPInt precision, OSQPInt max_active,
                          long long arena_bytes);

/* Queue a problem (setup + first solve).  x0 / y0: optional warm start.
 * Lengths: q and x0 hold n = P->n entries, l, u and y0 hold m = A->m; the
 * engine reads exactly that many (a shorter array is a caller error).  The
 * arrays (those of P and A included; the two OSQPCscMatrix structs are host
 * memory) may be host or device memory: device arrays are copied to the
 * host for validation and packing.
 * resident = 1 keeps it on the GPU after the solve for osqp_batch_resolve;
 * resident = 0 frees it once its result is polled.
 * Returns 0 with *id set, or OSQP_DATA_VALIDATION_ERROR. */
OSQPInt osqp_batch_add(OSQPBatch* batch, const OSQPCscMatrix* P, const OSQPFloat* q, const OSQPCscMatrix* A,
                       const OSQPFloat* l, const OSQPFloat* u, const OSQPFloat* x0, const OSQPFloat* y0,
                       OSQPInt resident, OSQPInt* id);

/* Queue count problems at once: the same as count osqp_batch_add calls in
 * order, with the host-side preparation (validation, CSR and row-group
 * images) spread over up to 8 threads.  P, q, A, l, u: arrays of count
 * pointers; x0 / y0: arrays of count pointers (entries may be null) or null.
 * ids[k] = the id of problem k, or -1 when it was rejected.  Returns 0 when
 * all were queued, else the error code of the first rejected problem.  No
 * counterpart in the reference (one osqp_setup per problem there). */
OSQPInt osqp_batch_add_many(OSQPBatch* batch, OSQPInt count, const OSQPCscMatrix* const* P, const OSQPFloat* const* q,
                            const OSQPCscMatrix* const* A, const OSQPFloat* const* l, const OSQPFloat* const* u,
                            const OSQPFloat* const* x0, const OSQPFloat* const* y0, OSQPInt resident,
                            OSQPInt* ids);

/* Queue a warm-started re-solve of a resident problem whose last solve has
 * ended and been polled; q_new (may be null, host or device memory)
 * replaces q first, with osqp_update_data_vec's semantics (src/osqp_api.c:
 * 1184-1257).  Matrix values staged by osqp_batch_update_data_mat since the
 * last solve are applied first.  Returns 0, -1 when the problem is unknown,
 * not resident or busy, or -2 when no staging memory is free for the staged
 * values (retry after polling other results). */
OSQPInt osqp_batch_resolve(OSQPBatch* batch, OSQPInt id, const OSQPFloat* q_new);

/* osqp_update_data_mat (src/osqp_api.c:1309-1409) for a resident problem
 * whose last solve has ended and been polled: new values of P (its upper
 * triangle, in the CSC order of osqp_batch_add) and / or A (CSC order), all
 * of them (index array null, count 0 or nnz) or those listed by the index
 * array.  Values and indices may be host or device memory.  The values are
 * staged; the next osqp_batch_resolve applies them on the GPU as the
 * reference does (unscale the data, write the values, scale again, rebuild
 * the preconditioner) before it solves, warm-started from the iterates.
 * Several calls before one resolve apply in call order.  Returns 0, 1 (P
 * count or indices invalid), 2 (A count or indices invalid), or -1 (unknown,
 * not resident or busy). */
OSQPInt osqp_batch_update_data_mat(OSQPBatch* batch, OSQPInt id, const OSQPFloat* Px_new,
                                   const OSQPInt* Px_new_idx, OSQPInt P_new_n, const OSQPFloat* Ax_new,
                                   const OSQPInt* Ax_new_idx, OSQPInt A_new_n);

/* Admits queued problems, waits until at least one bucket's epoch in flight
 * has ended (no wait when none is in flight), finishes every bucket whose
 * epoch has ended, and launches the next epoch of every idle bucket with work
 * without waiting for it.  *completed (may be null) = solves that ended in
 * the finished epochs.  Returns 0, or
 * OSQP_MEM_ALLOC_ERROR when a HIP call failed.  Loop with
 * `while (osqp_batch_pending(b)) osqp_batch_step(b, &done);`. */
OSQPInt osqp_batch_step(OSQPBatch* batch, OSQPInt* completed);

/* Ids of the solves that ended in the last osqp_batch_step (at most cap of
 * them written to ids); returns their number. */
OSQPInt osqp_batch_last_completed(const OSQPBatch* batch, OSQPInt* ids, OSQPInt cap);

/* osqp_batch_poll of count ids in one call: states[k] = 1 when problem k's
 * result was taken (infos[k] = its info; a non-resident problem is released),
 * 0 while it is queued or runs, 2 for a resident problem whose result was
 * taken before (idle), -1 for an unknown id; infos[k] is zeroed otherwise.
 * x / y (may be null): the x (n) and y (m) of the problems taken, one after
 * another in the order of ids.  Returns the number taken, or -1 for bad
 * arguments.  No counterpart in the reference. */
OSQPInt osqp_batch_poll_many(OSQPBatch* batch, OSQPInt count, const OSQPInt* ids, OSQPFloat* x, OSQPFloat* y,
                             OSQPBatchInfo* infos, OSQPInt* states);

/* 1 and the result when the problem's solve has ended; 0 while it runs; -1
 * for an unknown id.  x: n, y: m, prim_inf_cert: m, dual_inf_cert: n (any may
 * be null), as OSQPSolution holds them after osqp_solve.  A non-resident
 * problem is released by a successful poll. */
OSQPInt osqp_batch_poll(OSQPBatch* batch, OSQPInt id, OSQPFloat* x, OSQPFloat* y, OSQPFloat* prim_inf_cert,
                        OSQPFloat* dual_inf_cert, OSQPBatchInfo* info);

/* Drop a problem that is not running.  Returns 0 or -1. */
OSQPInt osqp_batch_release(OSQPBatch* batch, OSQPInt id);

/* Solves queued or running. */
OSQPInt osqp_batch_pending(const OSQPBatch* batch);

/* Tuned configuration of a size bucket (team = threads per workgroup,
 * teams_per_cu = resident workgroups per CU); 0, or -1 if not tuned yet. */
OSQPInt osqp_batch_tuning(const OSQPBatch* batch, OSQPInt bucket, OSQPInt* team, OSQPInt* teams_per_cu);

/* Pin a size bucket to one configuration (team in {64, 128, 256, 512},
 * teams_per_cu >= 1), which also stops the tuner for that bucket.  Results do
 * not change; this exists for tests and for callers that tune offline.
 * Returns 0 or -1. */
OSQPInt osqp_batch_set_team(OSQPBatch* batch, OSQPInt bucket, OSQPInt team, OSQPInt teams_per_cu);

/* Engine counters: epochs run, ADMM iterations, GPU time of the epochs (s:
 * the time during which at least one bucket's epoch ran). */
OSQPInt osqp_batch_counters(const OSQPBatch* batch, long long* epochs, long long* admm_iters, OSQPFloat* gpu_s);

OSQPInt osqp_batch_destroy(OSQPBatch* batch);

#ifdef __cplusplus
}
#endif

#endif /* OSQP_AMD_SERVE_H */
