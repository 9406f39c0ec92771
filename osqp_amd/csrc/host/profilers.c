/*
 * profilers.c — roctx section annotations and the Ctrl-C listener of the
 * host core (declarations and the section list in ../profilers.h).
 *
 * Annotations: src/profilers.c:4-28 gives each section a description and a
 * detail level; a section is emitted when its level is at most the solver's
 * profiler_level (src/profilers_roctx.c:23-42), as a roctx range named by its
 * description (rocprofiler-sdk roctx; `rocprofv3 --marker-trace` records
 * them).  The names are kept so that a rocprofv3 marker summary lines up with
 * the reference's profile tables (docs/advanced/profiling.rst:77-89).
 *
 * Interrupt: src/interrupt_unix.c:15-33 and src/osqp_api.c:690-737.
 */
#include <signal.h>
#include <string.h>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../profilers.h"
#include "../../../include/osqp_amd.h"
#include "../../../include/osqp_amd_algebra.h"

static const struct {
  const char* desc;
  int level;
} sections[OSQP_AMD_SEC_COUNT] = {
    {"Problem setup", 1},
    {"Problem data scaling", 1},
    {"Solving optimization problem", 1},
    {"ADMM iteration", 1},
    {"KKT system solve in ADMM iteration", 1},
    {"Vector updates in ADMM iteration", 1},
    {"Projection in ADMM iteration", 1},
    {"Solution polishing", 1},
    {"Initialize linear system solver", 2},
    {"Solve the linear system", 2},
    {"Symbolic factorization in direct solver", 2},
    {"Numeric factorization in direct solver", 2},
    {"Backsolve in direct solver", 2},
    {"Matrix-vector multiplication", 2},
};

static const struct {
  const char* desc;
  int level;
} events[OSQP_AMD_EVENT_COUNT] = {{"Rho update", 1}};

static int profiler_level;
static long long n_push, n_pop, n_mark;

void osqp_amd_profiler_level(int level) { profiler_level = level; }

int osqp_amd_profiler_enabled(osqp_amd_section sec) {
  return profiler_level > 0 && sections[sec].level <= profiler_level;
}

void osqp_amd_sec_push(osqp_amd_section sec) {
  if (!osqp_amd_profiler_enabled(sec)) return;
  roctxRangePushA(sections[sec].desc);
  n_push++;
}

void osqp_amd_sec_pop(osqp_amd_section sec) {
  if (!osqp_amd_profiler_enabled(sec)) return;
  osqp_amd_algebra_sync();
  roctxRangePop();
  n_pop++;
}

void osqp_amd_event_mark(osqp_amd_event ev) {
  if (profiler_level <= 0 || events[ev].level > profiler_level) return;
  roctxMarkA(events[ev].desc);
  n_mark++;
}

OSQPInt osqp_amd_profiler_counts(long long* pushes, long long* pops, long long* marks) {
  if (pushes) *pushes = n_push;
  if (pops) *pops = n_pop;
  if (marks) *marks = n_mark;
  return 0;
}

/* ---- Ctrl-C ---------------------------------------------------------------- */
static volatile sig_atomic_t interrupted;
static struct sigaction saved_action;

static void on_sigint(int sig) { interrupted = sig ? sig : -1; }

void osqp_amd_interrupt_listen(void) {
  struct sigaction act;
  memset(&act, 0, sizeof(act));
  interrupted = 0;
  sigemptyset(&act.sa_mask);
  act.sa_handler = on_sigint;
  sigaction(SIGINT, &act, &saved_action);
}

void osqp_amd_interrupt_restore(void) { sigaction(SIGINT, &saved_action, NULL); }

int osqp_amd_interrupted(void) { return interrupted; }
