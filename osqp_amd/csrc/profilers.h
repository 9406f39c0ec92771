/*
 * profilers.h — section annotations of the solver for rocprofv3, and the
 * Ctrl-C listener of osqp_solve.  Internal to libosqp_amd.so; usable from the
 * C host core and from the HIP sources.
 *
 * Sections, their detail levels and their names follow OSQP's profiler table
 * (src/profilers.c:4-28, enum include/private/profilers.h): a level-1 setting
 * annotates setup, scaling, the solve, each ADMM iteration and its KKT solve /
 * vector update / projection parts and polishing; level 2 adds the linear
 * system's init, solve and matrix-vector products.  They are emitted as roctx
 * ranges (rocprofiler-sdk), which `rocprofv3 --marker-trace` records, as the
 * reference's roctx backend does (src/profilers_roctx.c:44-68).
 */
#ifndef OSQP_AMD_PROFILERS_H
#define OSQP_AMD_PROFILERS_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  OSQP_AMD_SEC_SETUP = 0,
  OSQP_AMD_SEC_SCALE,
  OSQP_AMD_SEC_OPT_SOLVE,
  OSQP_AMD_SEC_ADMM_ITER,
  OSQP_AMD_SEC_ADMM_KKT_SOLVE,
  OSQP_AMD_SEC_ADMM_UPDATE,
  OSQP_AMD_SEC_ADMM_PROJ,
  OSQP_AMD_SEC_POLISH,
  OSQP_AMD_SEC_LINSYS_INIT,
  OSQP_AMD_SEC_LINSYS_SOLVE,
  OSQP_AMD_SEC_LINSYS_SYM_FAC,
  OSQP_AMD_SEC_LINSYS_NUM_FAC,
  OSQP_AMD_SEC_LINSYS_BACKSOLVE,
  OSQP_AMD_SEC_LINSYS_MVM,
  OSQP_AMD_SEC_COUNT
} osqp_amd_section;

typedef enum { OSQP_AMD_EVENT_RHO_UPDATE = 0, OSQP_AMD_EVENT_COUNT } osqp_amd_event;

/* Sets the detail level (0: off, 1, 2), process-wide like the reference's
 * profiler tables. */
void osqp_amd_profiler_level(int level);
int  osqp_amd_profiler_enabled(osqp_amd_section sec);
void osqp_amd_sec_push(osqp_amd_section sec);
/* Ends the section.  The backend works asynchronously, so the pop of an
 * enabled section first waits for the GPU work queued inside it: the range
 * then covers that work, as the reference's synchronous CUDA calls do. */
void osqp_amd_sec_pop(osqp_amd_section sec);
void osqp_amd_event_mark(osqp_amd_event ev);

/* Ctrl-C during a solve (src/interrupt_unix.c:15-33): the listener replaces
 * the SIGINT action for the duration of osqp_solve and restores it after. */
void osqp_amd_interrupt_listen(void);
void osqp_amd_interrupt_restore(void);
int  osqp_amd_interrupted(void);

#ifdef __cplusplus
}
#endif

#endif
