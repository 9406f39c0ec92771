/*
 * osqp_oracle.c — CPU FP64 restatement of OSQP's ADMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the
 * MI355X engine and the CPU leg of the benchmark.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as
 * the checker; the product library never links, loads or calls it.
 *
 * What it restates (reference osqp/osqp, file:line):
 *   ADMM driver            src/osqp_api.c:636-1038 (loop 698-903)
 *   setup / updates        src/osqp_api.c:332-631, 1184-1464, 1474-1530
 *   ADMM steps, residuals, termination, infeasibility, rho adaptation,
 *   solution store         src/auxil.c:14-945
 *   Ruiz scaling           src/scaling.c:48-208
 *   vector semantics       algebra/builtin/vector.c:315-953
 *   CSC kernels            algebra/_common/csc_math.c:26-360
 *   KKT assembly           algebra/_common/kkt.c:293-342
 *   direct KKT solve       algebra/_common/lin_sys/qdldl/qdldl_interface.c:85-134, 393-462
 *   reduced-KKT operator   algebra/_common/reduced_kkt.c:9-65
 *   PCG + tolerance        algebra/cuda/lin_sys/indirect/cuda_pcg.cu:113-266,
 *                          algebra/cuda/lin_sys/indirect/cuda_pcg_interface.cu:32-92,229-273
 *
 * The reference's direct solver is QDLDL v0.1.8 with SuiteSparse AMD ordering.
 * QDLDL is fetched at configure time and is absent from this image, so the
 * factorisation here is an independent implementation of the published
 * up-looking LDL^T algorithm (T. A. Davis, "Algorithm 849: a concise sparse
 * Cholesky factorization package", ACM TOMS 31(4), 2005) over an exact
 * minimum-degree ordering.  Any exact factorisation of the same quasi-definite
 * KKT matrix yields the same ADMM iterates up to rounding; the ordering and
 * therefore the last bits differ from QDLDL+AMD ("parity unpinned" at the
 * level of L and D values, pinned at solver level by the reference fixtures).
 *
 * Build: plain C99, gcc -O2 -fPIC -shared (oracle/Makefile).  FP64 throughout,
 * no FMA contraction (-ffp-contract=off) so that elementwise results are the
 * correctly rounded IEEE values the GPU kernels are compared against bit-wise.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../include/osqp_amd.h"

#define c_max(a, b) (((a) > (b)) ? (a) : (b))
#define c_min(a, b) (((a) < (b)) ? (a) : (b))
#define c_absval(x) (((x) < 0) ? -(x) : (x))

/* ========================================================================
 * Plain CSC helpers (algebra/_common/csc_math.c)
 * ====================================================================== */

typedef struct {
  int m, n;
  int *p, *i;
  double *x;
} ocsc;

static ocsc* ocsc_copy(const OSQPCscMatrix* A) {
  ocsc* M = (ocsc*)calloc(1, sizeof(ocsc));
  int nnz = A->p[A->n];
  M->m = A->m;
  M->n = A->n;
  M->p = (int*)malloc(sizeof(int) * (A->n + 1));
  M->i = (int*)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
  M->x = (double*)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
  memcpy(M->p, A->p, sizeof(int) * (A->n + 1));
  if (nnz) {
    memcpy(M->i, A->i, sizeof(int) * nnz);
    memcpy(M->x, A->x, sizeof(double) * nnz);
  }
  return M;
}

static void ocsc_free(ocsc* M) {
  if (!M) return;
  free(M->p);
  free(M->i);
  free(M->x);
  free(M);
}

static int ocsc_nnz(const ocsc* M) { return M->p[M->n]; }

/* y = alpha*A*x + beta*y  — csc_math.c:169-212 */
static void csc_Axpy(const ocsc* A, const double* x, double* y, double alpha, double beta) {
  int i, j;
  if (beta == 0) for (i = 0; i < A->m; i++) y[i] = 0.0;
  else if (beta == 1) ;
  else if (beta == -1) for (i = 0; i < A->m; i++) y[i] = -y[i];
  else for (i = 0; i < A->m; i++) y[i] *= beta;
  if (A->p[A->n] == 0 || alpha == 0.0) return;
  if (alpha == -1) {
    for (j = 0; j < A->n; j++) for (i = A->p[j]; i < A->p[j + 1]; i++) y[A->i[i]] -= A->x[i] * x[j];
  } else if (alpha == 1) {
    for (j = 0; j < A->n; j++) for (i = A->p[j]; i < A->p[j + 1]; i++) y[A->i[i]] += A->x[i] * x[j];
  } else {
    for (j = 0; j < A->n; j++) for (i = A->p[j]; i < A->p[j + 1]; i++) y[A->i[i]] += alpha * A->x[i] * x[j];
  }
}

/* y = alpha*A'*x + beta*y — csc_math.c:216-255 */
static void csc_Atxpy(const ocsc* A, const double* x, double* y, double alpha, double beta) {
  int j, k;
  if (beta == 0) for (j = 0; j < A->n; j++) y[j] = 0.0;
  else if (beta == 1) ;
  else if (beta == -1) for (j = 0; j < A->n; j++) y[j] = -y[j];
  else for (j = 0; j < A->n; j++) y[j] *= beta;
  if (A->p[A->n] == 0 || alpha == 0.0) return;
  if (alpha == -1) {
    for (j = 0; j < A->n; j++) for (k = A->p[j]; k < A->p[j + 1]; k++) y[j] -= A->x[k] * x[A->i[k]];
  } else if (alpha == 1) {
    for (j = 0; j < A->n; j++) for (k = A->p[j]; k < A->p[j + 1]; k++) y[j] += A->x[k] * x[A->i[k]];
  } else {
    for (j = 0; j < A->n; j++) for (k = A->p[j]; k < A->p[j + 1]; k++) y[j] += alpha * A->x[k] * x[A->i[k]];
  }
}

/* y = alpha*P*x + beta*y with only triu(P) stored — csc_math.c:114-166 */
static void csc_Axpy_sym_triu(const ocsc* A, const double* x, double* y, double alpha, double beta) {
  int i, j;
  if (beta == 0) for (i = 0; i < A->m; i++) y[i] = 0.0;
  else if (beta == 1) ;
  else if (beta == -1) for (i = 0; i < A->m; i++) y[i] = -y[i];
  else for (i = 0; i < A->m; i++) y[i] *= beta;
  if (A->p[A->n] == 0 || alpha == 0.0) return;
  if (alpha == -1) {
    for (j = 0; j < A->n; j++) for (i = A->p[j]; i < A->p[j + 1]; i++) {
      y[A->i[i]] -= A->x[i] * x[j];
      if (A->i[i] != j) y[j] -= A->x[i] * x[A->i[i]];
    }
  } else if (alpha == 1) {
    for (j = 0; j < A->n; j++) for (i = A->p[j]; i < A->p[j + 1]; i++) {
      y[A->i[i]] += A->x[i] * x[j];
      if (A->i[i] != j) y[j] += A->x[i] * x[A->i[i]];
    }
  } else {
    for (j = 0; j < A->n; j++) for (i = A->p[j]; i < A->p[j + 1]; i++) {
      y[A->i[i]] += alpha * A->x[i] * x[j];
      if (A->i[i] != j) y[j] += alpha * A->x[i] * x[A->i[i]];
    }
  }
}

/* csc_math.c:292-308: max |M_ij| over the STORED entries of column j (for the
 * triu-stored P of the builtin backend this is the triu column only) */
static void csc_col_norm_inf(const ocsc* M, double* E) {
  int j, k;
  for (j = 0; j < M->n; j++) {
    E[j] = 0.0;
    for (k = M->p[j]; k < M->p[j + 1]; k++) E[j] = c_max(c_absval(M->x[k]), E[j]);
  }
}

/* csc_math.c:312-331 */
static void csc_row_norm_inf(const ocsc* M, double* E) {
  int j, k;
  for (j = 0; j < M->m; j++) E[j] = 0.0;
  for (j = 0; j < M->n; j++)
    for (k = M->p[j]; k < M->p[j + 1]; k++) E[M->i[k]] = c_max(c_absval(M->x[k]), E[M->i[k]]);
}

/* csc_math.c:335-360: row norms of the full symmetric matrix from its triu */
static void csc_row_norm_inf_sym_triu(const ocsc* M, double* E) {
  int j, k, i;
  double a;
  for (j = 0; j < M->m; j++) E[j] = 0.0;
  for (j = 0; j < M->n; j++)
    for (k = M->p[j]; k < M->p[j + 1]; k++) {
      i = M->i[k];
      a = c_absval(M->x[k]);
      E[i] = c_max(a, E[i]);
      if (i != j) E[j] = c_max(a, E[j]);
    }
}

static void csc_lmult_diag(ocsc* A, const double* d) {
  int j, k;
  for (j = 0; j < A->n; j++) for (k = A->p[j]; k < A->p[j + 1]; k++) A->x[k] *= d[A->i[k]];
}

static void csc_rmult_diag(ocsc* A, const double* d) {
  int j, k;
  for (j = 0; j < A->n; j++) for (k = A->p[j]; k < A->p[j + 1]; k++) A->x[k] *= d[j];
}

static void csc_scale(ocsc* A, double sc) {
  int k;
  for (k = 0; k < ocsc_nnz(A); k++) A->x[k] *= sc;
}

/* ========================================================================
 * Exported kernel-level restatements (used by the algebra-boundary tests)
 * ====================================================================== */

static ocsc ocsc_view(int m, int n, const int* p, const int* i, const double* x) {
  ocsc M;
  M.m = m; M.n = n; M.p = (int*)p; M.i = (int*)i; M.x = (double*)x;
  return M;
}

void oracle_csc_Axpy(int m, int n, const int* p, const int* i, const double* x,
                     const double* v, double* y, double alpha, double beta) {
  ocsc M = ocsc_view(m, n, p, i, x);
  csc_Axpy(&M, v, y, alpha, beta);
}

void oracle_csc_Atxpy(int m, int n, const int* p, const int* i, const double* x,
                      const double* v, double* y, double alpha, double beta) {
  ocsc M = ocsc_view(m, n, p, i, x);
  csc_Atxpy(&M, v, y, alpha, beta);
}

void oracle_csc_Axpy_sym_triu(int n, const int* p, const int* i, const double* x,
                              const double* v, double* y, double alpha, double beta) {
  ocsc M = ocsc_view(n, n, p, i, x);
  csc_Axpy_sym_triu(&M, v, y, alpha, beta);
}

void oracle_csc_col_norm_inf(int m, int n, const int* p, const int* i, const double* x, double* E) {
  ocsc M = ocsc_view(m, n, p, i, x);
  csc_col_norm_inf(&M, E);
}

void oracle_csc_row_norm_inf(int m, int n, const int* p, const int* i, const double* x, double* E) {
  ocsc M = ocsc_view(m, n, p, i, x);
  csc_row_norm_inf(&M, E);
}

void oracle_csc_row_norm_inf_sym_triu(int n, const int* p, const int* i, const double* x, double* E) {
  ocsc M = ocsc_view(n, n, p, i, x);
  csc_row_norm_inf_sym_triu(&M, E);
}

/* d = diag(A' diag(D) A) — csc_math.c:94-111 */
void oracle_csc_AtDA_extract_diag(int m, int n, const int* p, const int* i, const double* x,
                                  const double* D, double* d) {
  int j, k;
  (void)m;
  for (j = 0; j < n; j++) {
    d[j] = 0;
    for (k = p[j]; k < p[j + 1]; k++) d[j] += x[k] * x[k] * D[i[k]];
  }
}

/* Structure oracle: CSR of an m x n CSC matrix by a stable counting sort over
 * rows (entries of a row appear in increasing column order).  csr_pos[k] is the
 * CSR slot of CSC entry k.  This is the ordering csc_utils.c:247-286
 * (triplet_to_csr) produces for column-sorted input. */
void oracle_csc_to_csr(int m, int n, const int* p, const int* i,
                       int* rp, int* ci, int* csr_pos) {
  int j, k, r;
  int* cnt = (int*)calloc((size_t)m + 1, sizeof(int));
  for (k = 0; k < p[n]; k++) cnt[i[k] + 1]++;
  for (r = 0; r < m; r++) cnt[r + 1] += cnt[r];
  memcpy(rp, cnt, sizeof(int) * ((size_t)m + 1));
  for (j = 0; j < n; j++)
    for (k = p[j]; k < p[j + 1]; k++) {
      int pos = cnt[i[k]]++;
      ci[pos] = j;
      csr_pos[k] = pos;
    }
  free(cnt);
}

/* Structure oracle: full symmetric CSR (rows sorted by column) of an n x n
 * matrix given by its upper triangle in CSC.  Entry k of the triu CSC lands at
 * full_pos[k] (its own position, row = stored row) and mirror_pos[k] (the
 * transposed position, -1 on the diagonal).  Matches the content of
 * csc_utils.c triu_to_csc; the CUDA backend's triu->full conversion
 * (cuda_csr.cu:489-549) differs only in adding explicit zero diagonals. */
void oracle_triu_to_full_csr(int n, const int* p, const int* i,
                             int* rp, int* ci, int* full_pos, int* mirror_pos) {
  int j, k, r;
  int nnz = p[n];
  int* cnt = (int*)calloc((size_t)n + 1, sizeof(int));
  /* row counts of the full matrix */
  for (j = 0; j < n; j++)
    for (k = p[j]; k < p[j + 1]; k++) {
      cnt[i[k] + 1]++;
      if (i[k] != j) cnt[j + 1]++;
    }
  for (r = 0; r < n; r++) cnt[r + 1] += cnt[r];
  memcpy(rp, cnt, sizeof(int) * ((size_t)n + 1));
  /* Row r of the full matrix holds, in increasing column order:
   *   (r, c) for c < r : mirrored triu entries (c, r), i.e. column r of the
   *                      triu CSC with row c < r  -> scanned in CSC order
   *   (r, c) for c >= r: triu entries of row r, found in columns c >= r.
   * A single pass over columns j = 0..n-1 visits the targets of each row in
   * increasing column order for both kinds, provided the mirrored entries of
   * row j (which come from column j) are placed before row j receives any
   * c >= j entry.  Column j's entries (i<j) go to row i at column j (i<j, so
   * j is after all previously placed columns of row i) and mirrored to row j
   * at columns i < j in increasing i (CSC rows sorted). */
  for (j = 0; j < n; j++) {
    for (k = p[j]; k < p[j + 1]; k++) {
      int row = i[k];
      if (row == j) continue;
      /* mirrored (j, row) into row j, column row < j */
      int pos = cnt[j]++;
      ci[pos] = row;
      mirror_pos[k] = pos;
    }
    for (k = p[j]; k < p[j + 1]; k++) {
      int row = i[k];
      int pos = cnt[row]++;
      ci[pos] = j;
      full_pos[k] = pos;
      if (row == j) mirror_pos[k] = -1;
    }
  }
  (void)nnz;
  free(cnt);
}

/* ========================================================================
 * Minimum-degree ordering on the symmetric pattern of an upper CSC matrix
 * (exact external degree on the elimination graph, LIFO tie-breaking).
 * ====================================================================== */

static void order_min_degree(int N, const int* Kp, const int* Ki, int* perm) {
  int v, k, s;
  int* len = (int*)calloc((size_t)N, sizeof(int));
  int* cap = (int*)calloc((size_t)N, sizeof(int));
  int** adj = (int**)calloc((size_t)N, sizeof(int*));
  int* deg = (int*)calloc((size_t)N, sizeof(int));
  int* head = (int*)malloc(sizeof(int) * ((size_t)N + 1));
  int* next = (int*)malloc(sizeof(int) * (size_t)N);
  int* prev = (int*)malloc(sizeof(int) * (size_t)N);
  int* mark = (int*)calloc((size_t)N, sizeof(int));
  char* alive = (char*)malloc((size_t)N);
  int* nb = (int*)malloc(sizeof(int) * (size_t)N);
  int* tmp = (int*)malloc(sizeof(int) * (size_t)N);
  int stamp = 0, mindeg = 0;

  for (v = 0; v < N; v++) {
    for (k = Kp[v]; k < Kp[v + 1]; k++) {
      int r = Ki[k];
      if (r == v) continue;
      len[r]++;
      len[v]++;
    }
  }
  for (v = 0; v < N; v++) {
    cap[v] = len[v] > 4 ? len[v] : 4;
    adj[v] = (int*)malloc(sizeof(int) * (size_t)cap[v]);
    len[v] = 0;
  }
  for (v = 0; v < N; v++) {
    for (k = Kp[v]; k < Kp[v + 1]; k++) {
      int r = Ki[k];
      if (r == v) continue;
      adj[r][len[r]++] = v;
      adj[v][len[v]++] = r;
    }
  }
  /* dedupe (input has no duplicates, but be safe) */
  for (v = 0; v < N; v++) {
    int w, nl = 0;
    stamp++;
    for (w = 0; w < len[v]; w++) {
      int u = adj[v][w];
      if (mark[u] != stamp) { mark[u] = stamp; adj[v][nl++] = u; }
    }
    len[v] = nl;
    deg[v] = nl;
  }
  for (v = 0; v <= N; v++) head[v] = -1;
  for (v = N - 1; v >= 0; v--) {
    next[v] = head[deg[v]];
    prev[v] = -1;
    if (head[deg[v]] >= 0) prev[head[deg[v]]] = v;
    head[deg[v]] = v;
    alive[v] = 1;
  }

  for (s = 0; s < N; s++) {
    int nnb = 0, w;
    while (head[mindeg] < 0) mindeg++;
    v = head[mindeg];
    /* unlink v */
    head[mindeg] = next[v];
    if (next[v] >= 0) prev[next[v]] = -1;
    alive[v] = 0;
    perm[s] = v;
    for (w = 0; w < len[v]; w++) if (alive[adj[v][w]]) nb[nnb++] = adj[v][w];
    for (w = 0; w < nnb; w++) {
      int u = nb[w], t, nl = 0;
      stamp++;
      for (t = 0; t < len[u]; t++) {
        int x = adj[u][t];
        if (alive[x] && mark[x] != stamp) { mark[x] = stamp; tmp[nl++] = x; }
      }
      for (t = 0; t < nnb; t++) {
        int x = nb[t];
        if (x != u && mark[x] != stamp) { mark[x] = stamp; tmp[nl++] = x; }
      }
      if (nl > cap[u]) {
        cap[u] = nl + nl / 2 + 4;
        adj[u] = (int*)realloc(adj[u], sizeof(int) * (size_t)cap[u]);
      }
      memcpy(adj[u], tmp, sizeof(int) * (size_t)nl);
      len[u] = nl;
      /* move u to bucket nl */
      if (prev[u] >= 0) next[prev[u]] = next[u];
      else head[deg[u]] = next[u];
      if (next[u] >= 0) prev[next[u]] = prev[u];
      deg[u] = nl;
      next[u] = head[nl];
      prev[u] = -1;
      if (head[nl] >= 0) prev[head[nl]] = u;
      head[nl] = u;
    }
    free(adj[v]);
    adj[v] = NULL;
    mindeg = mindeg > 0 ? mindeg - 1 : 0;
  }
  free(len); free(cap); free(adj); free(deg); free(head); free(next); free(prev);
  free(mark); free(alive); free(nb); free(tmp);
}

/* ========================================================================
 * Up-looking sparse LDL^T (Davis 2005) on an upper-triangular CSC matrix
 * ====================================================================== */

typedef struct {
  int N;
  int *Lp, *Li, *Lnz, *parent, *flag, *pattern;
  double *Lx, *D, *Dinv, *Y;
} oldl;

static oldl* ldl_symbolic(int N, const int* Ap, const int* Ai) {
  int k, p;
  oldl* F = (oldl*)calloc(1, sizeof(oldl));
  F->N = N;
  F->Lp = (int*)malloc(sizeof(int) * ((size_t)N + 1));
  F->Lnz = (int*)malloc(sizeof(int) * (size_t)N);
  F->parent = (int*)malloc(sizeof(int) * (size_t)N);
  F->flag = (int*)malloc(sizeof(int) * (size_t)N);
  F->pattern = (int*)malloc(sizeof(int) * (size_t)N);
  F->D = (double*)malloc(sizeof(double) * (size_t)N);
  F->Dinv = (double*)malloc(sizeof(double) * (size_t)N);
  F->Y = (double*)calloc((size_t)N, sizeof(double));
  for (k = 0; k < N; k++) {
    F->parent[k] = -1;
    F->flag[k] = k;
    F->Lnz[k] = 0;
    for (p = Ap[k]; p < Ap[k + 1]; p++) {
      int i = Ai[p];
      if (i >= k) continue;
      for (; F->flag[i] != k; i = F->parent[i]) {
        if (F->parent[i] == -1) F->parent[i] = k;
        F->Lnz[i]++;
        F->flag[i] = k;
      }
    }
  }
  F->Lp[0] = 0;
  for (k = 0; k < N; k++) F->Lp[k + 1] = F->Lp[k] + F->Lnz[k];
  F->Li = (int*)malloc(sizeof(int) * (size_t)(F->Lp[N] > 0 ? F->Lp[N] : 1));
  F->Lx = (double*)malloc(sizeof(double) * (size_t)(F->Lp[N] > 0 ? F->Lp[N] : 1));
  return F;
}

/* returns the number of positive pivots, or -1 on a zero pivot */
static int ldl_numeric(oldl* F, const int* Ap, const int* Ai, const double* Ax) {
  int k, p, N = F->N, npos = 0;
  for (k = 0; k < N; k++) {
    int top = N;
    double yi, lki;
    F->Y[k] = 0.0;
    F->flag[k] = k;
    F->Lnz[k] = 0;
    for (p = Ap[k]; p < Ap[k + 1]; p++) {
      int i = Ai[p], len;
      F->Y[i] += Ax[p];
      for (len = 0; F->flag[i] != k; i = F->parent[i]) {
        F->pattern[len++] = i;
        F->flag[i] = k;
      }
      while (len > 0) F->pattern[--top] = F->pattern[--len];
    }
    F->D[k] = F->Y[k];
    F->Y[k] = 0.0;
    for (; top < N; top++) {
      int i = F->pattern[top], p2 = F->Lp[i] + F->Lnz[i];
      yi = F->Y[i];
      F->Y[i] = 0.0;
      for (p = F->Lp[i]; p < p2; p++) F->Y[F->Li[p]] -= F->Lx[p] * yi;
      lki = yi / F->D[i];
      F->D[k] -= lki * yi;
      F->Li[p2] = k;
      F->Lx[p2] = lki;
      F->Lnz[i]++;
    }
    if (F->D[k] == 0.0) return -1;
    if (F->D[k] > 0.0) npos++;
    F->Dinv[k] = 1.0 / F->D[k];
  }
  return npos;
}

static void ldl_solve(const oldl* F, double* x) {
  int j, p, N = F->N;
  for (j = 0; j < N; j++)
    for (p = F->Lp[j]; p < F->Lp[j + 1]; p++) x[F->Li[p]] -= F->Lx[p] * x[j];
  for (j = 0; j < N; j++) x[j] *= F->Dinv[j];
  for (j = N - 1; j >= 0; j--)
    for (p = F->Lp[j]; p < F->Lp[j + 1]; p++) x[j] -= F->Lx[p] * x[F->Li[p]];
}

static void ldl_free(oldl* F) {
  if (!F) return;
  free(F->Lp); free(F->Li); free(F->Lnz); free(F->parent); free(F->flag); free(F->pattern);
  free(F->Lx); free(F->D); free(F->Dinv); free(F->Y);
  free(F);
}

/* ========================================================================
 * Linear-system solvers behind the linsys "vtable" (include/private/types.h:243-279)
 * ====================================================================== */

typedef struct oracle_linsys oracle_linsys;
struct oracle_linsys {
  int type;   /* OSQP_DIRECT_SOLVER or OSQP_INDIRECT_SOLVER */
  int n, m;
  double sigma, rho_sc;
  const ocsc *P, *A;          /* borrowed (scaled) problem data */
  const double* rho_vec;      /* borrowed, NULL when rho is scalar */
  /* direct */
  int N, *Kp, *Ki, *perm, *pinv, *rho_pos, *sigma_pos;
  double *Kx, *bp, *sol;
  oldl* F;
  /* indirect (reduced KKT PCG, FP64) */
  int max_iter, reduction_threshold, zero_pcg_iters, precond_type;
  double tol_fraction, reduction_factor, eps_prev;
  const double *scaled_prim_res, *scaled_dual_res;
  double *px, *pp, *pKp, *py, *pr, *rhs, *pz, *precond_inv;
  long total_pcg_iters;
};

/* -- direct: KKT [P+sigma I, A'; A, -diag(1/rho)] (kkt.c:293-342) -------- */

static int kkt_assemble_factor(oracle_linsys* s, int first) {
  int n = s->n, m = s->m, N = n + m, j, k, i;
  if (first) {
    /* unpermuted upper CSC */
    int* Up = (int*)calloc((size_t)N + 1, sizeof(int));
    int *Ui, nnzU, *cnt, *rowcnt;
    double* Ux;
    int* srcP;   /* per KKT entry: code for value source */
    int* diagP = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (j = 0; j < n; j++) {
      diagP[j] = -1;
      for (k = s->P->p[j]; k < s->P->p[j + 1]; k++) if (s->P->i[k] == j) diagP[j] = k;
      Up[j + 1] = (s->P->p[j + 1] - s->P->p[j]) + (diagP[j] < 0 ? 1 : 0);
    }
    rowcnt = (int*)calloc((size_t)(m > 0 ? m : 1), sizeof(int));
    for (k = 0; k < s->A->p[n]; k++) rowcnt[s->A->i[k]]++;
    for (i = 0; i < m; i++) Up[n + i + 1] = rowcnt[i] + 1;
    for (j = 0; j < N; j++) Up[j + 1] += Up[j];
    nnzU = Up[N];
    Ui = (int*)malloc(sizeof(int) * (size_t)nnzU);
    Ux = (double*)malloc(sizeof(double) * (size_t)nnzU);
    srcP = (int*)malloc(sizeof(int) * (size_t)nnzU);
    cnt = (int*)malloc(sizeof(int) * (size_t)N);
    for (j = 0; j < N; j++) cnt[j] = Up[j];
    s->sigma_pos = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    s->rho_pos = (int*)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    for (j = 0; j < n; j++) {
      for (k = s->P->p[j]; k < s->P->p[j + 1]; k++) {
        int pos = cnt[j]++;
        Ui[pos] = s->P->i[k];
        Ux[pos] = s->P->x[k] + (s->P->i[k] == j ? s->sigma : 0.0);
        if (s->P->i[k] == j) s->sigma_pos[j] = pos;
      }
      if (diagP[j] < 0) {
        int pos = cnt[j]++;
        Ui[pos] = j;
        Ux[pos] = s->sigma;
        s->sigma_pos[j] = pos;
      }
    }
    for (j = 0; j < n; j++)
      for (k = s->A->p[j]; k < s->A->p[j + 1]; k++) {
        int col = n + s->A->i[k];
        int pos = cnt[col]++;
        Ui[pos] = j;
        Ux[pos] = s->A->x[k];
      }
    for (i = 0; i < m; i++) {
      int pos = cnt[n + i]++;
      Ui[pos] = n + i;
      Ux[pos] = s->rho_vec ? -1.0 / s->rho_vec[i] : -1.0 / s->rho_sc;
      s->rho_pos[i] = pos;
    }
    /* ordering */
    s->N = N;
    s->perm = (int*)malloc(sizeof(int) * (size_t)N);
    s->pinv = (int*)malloc(sizeof(int) * (size_t)N);
    order_min_degree(N, Up, Ui, s->perm);
    for (j = 0; j < N; j++) s->pinv[s->perm[j]] = j;
    /* permuted upper CSC; keep map old entry -> new entry */
    {
      int* Cp = (int*)calloc((size_t)N + 1, sizeof(int));
      int* map = (int*)malloc(sizeof(int) * (size_t)nnzU);
      for (j = 0; j < N; j++)
        for (k = Up[j]; k < Up[j + 1]; k++) {
          int a = s->pinv[Ui[k]], b = s->pinv[j];
          Cp[(a > b ? a : b) + 1]++;
        }
      for (j = 0; j < N; j++) Cp[j + 1] += Cp[j];
      for (j = 0; j < N; j++) cnt[j] = Cp[j];
      s->Ki = (int*)malloc(sizeof(int) * (size_t)nnzU);
      s->Kx = (double*)malloc(sizeof(double) * (size_t)nnzU);
      for (j = 0; j < N; j++)
        for (k = Up[j]; k < Up[j + 1]; k++) {
          int a = s->pinv[Ui[k]], b = s->pinv[j];
          int col = a > b ? a : b, row = a > b ? b : a;
          int pos = cnt[col]++;
          s->Ki[pos] = row;
          s->Kx[pos] = Ux[k];
          map[k] = pos;
        }
      s->Kp = Cp;
      for (j = 0; j < n; j++) s->sigma_pos[j] = map[s->sigma_pos[j]];
      for (i = 0; i < m; i++) s->rho_pos[i] = map[s->rho_pos[i]];
      free(map);
    }
    free(Up); free(Ui); free(Ux); free(srcP); free(cnt); free(rowcnt); free(diagP);
    s->F = ldl_symbolic(N, s->Kp, s->Ki);
    s->bp = (double*)malloc(sizeof(double) * (size_t)N);
    s->sol = (double*)malloc(sizeof(double) * (size_t)N);
  }
  {
    int npos = ldl_numeric(s->F, s->Kp, s->Ki, s->Kx);
    /* qdldl_interface.c:125-131: zero pivot, or fewer than n positive pivots */
    if (npos < 0) return OSQP_NONCVX_ERROR;
    if (npos < n) return OSQP_NONCVX_ERROR;
  }
  return 0;
}

/* refresh the KKT values from the (possibly rescaled) P, A and rho;
 * mirrors update_KKT_P / update_KKT_A / update_KKT_param2 (kkt.c:349-422) */
static int kkt_refresh(oracle_linsys* s) {
  int n = s->n, m = s->m, j, k, i;
  /* rebuild values by re-walking the unpermuted layout */
  int N = n + m;
  int* cnt = (int*)malloc(sizeof(int) * (size_t)N);
  (void)cnt;
  free(cnt);
  /* simplest exact approach: rebuild from scratch keeping the ordering */
  {
    int* Up = (int*)calloc((size_t)N + 1, sizeof(int));
    int* rowcnt = (int*)calloc((size_t)(m > 0 ? m : 1), sizeof(int));
    int *Ui, *cc, nnzU;
    double* Ux;
    for (j = 0; j < n; j++) {
      int hasd = 0;
      for (k = s->P->p[j]; k < s->P->p[j + 1]; k++) if (s->P->i[k] == j) hasd = 1;
      Up[j + 1] = (s->P->p[j + 1] - s->P->p[j]) + (hasd ? 0 : 1);
    }
    for (k = 0; k < s->A->p[n]; k++) rowcnt[s->A->i[k]]++;
    for (i = 0; i < m; i++) Up[n + i + 1] = rowcnt[i] + 1;
    for (j = 0; j < N; j++) Up[j + 1] += Up[j];
    nnzU = Up[N];
    Ui = (int*)malloc(sizeof(int) * (size_t)nnzU);
    Ux = (double*)malloc(sizeof(double) * (size_t)nnzU);
    cc = (int*)malloc(sizeof(int) * (size_t)N);
    for (j = 0; j < N; j++) cc[j] = Up[j];
    for (j = 0; j < n; j++) {
      int hasd = 0;
      for (k = s->P->p[j]; k < s->P->p[j + 1]; k++) {
        int pos = cc[j]++;
        Ui[pos] = s->P->i[k];
        Ux[pos] = s->P->x[k] + (s->P->i[k] == j ? s->sigma : 0.0);
        if (s->P->i[k] == j) hasd = 1;
      }
      if (!hasd) { int pos = cc[j]++; Ui[pos] = j; Ux[pos] = s->sigma; }
    }
    for (j = 0; j < n; j++)
      for (k = s->A->p[j]; k < s->A->p[j + 1]; k++) {
        int pos = cc[n + s->A->i[k]]++;
        Ui[pos] = j;
        Ux[pos] = s->A->x[k];
      }
    for (i = 0; i < m; i++) {
      int pos = cc[n + i]++;
      Ui[pos] = n + i;
      Ux[pos] = s->rho_vec ? -1.0 / s->rho_vec[i] : -1.0 / s->rho_sc;
    }
    /* same permutation walk as at assembly: entries land at the same slots */
    for (j = 0; j < N; j++) cc[j] = s->Kp[j];
    for (j = 0; j < N; j++)
      for (k = Up[j]; k < Up[j + 1]; k++) {
        int a = s->pinv[Ui[k]], b = s->pinv[j];
        int col = a > b ? a : b;
        int pos = cc[col]++;
        s->Kx[pos] = Ux[k];
      }
    free(Up); free(rowcnt); free(Ui); free(Ux); free(cc);
  }
  return kkt_assemble_factor(s, 0);
}

/* qdldl_interface.c:418-462 */
static void direct_solve(oracle_linsys* s, double* b) {
  int j, n = s->n, m = s->m, N = s->N;
  for (j = 0; j < N; j++) s->bp[j] = b[s->perm[j]];
  ldl_solve(s->F, s->bp);
  for (j = 0; j < N; j++) s->sol[s->perm[j]] = s->bp[j];
  for (j = 0; j < n; j++) b[j] = s->sol[j];
  for (j = 0; j < m; j++)
    b[j + n] += (s->rho_vec ? 1.0 / s->rho_vec[j] : 1.0 / s->rho_sc) * s->sol[j + n];
}

/* -- indirect: FP64 PCG on K = P + sigma I + A' R A (reduced_kkt.c) ------- */

static double vdot(const double* a, const double* b, int n) {
  double s = 0.0;
  int i;
  for (i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

static double vnorm_inf(const double* a, int n) {
  double s = 0.0;
  int i;
  for (i = 0; i < n; i++) { double v = c_absval(a[i]); if (v > s) s = v; }
  return s;
}

/* y = (P + sigma I + A' R A) x — reduced_kkt.c:9-24, cuda_pcg.cu:50-106 */
static void pcg_matvec(oracle_linsys* s, const double* x, double* y) {
  int i;
  for (i = 0; i < s->n; i++) y[i] = s->sigma * x[i];
  csc_Axpy_sym_triu(s->P, x, y, 1.0, 1.0);
  if (s->m == 0) return;
  csc_Axpy(s->A, x, s->pz, 1.0, 0.0);
  for (i = 0; i < s->m; i++) s->pz[i] *= (s->rho_vec ? s->rho_vec[i] : s->rho_sc);
  csc_Atxpy(s->A, s->pz, y, 1.0, 1.0);
}

/* cuda_pcg.cu:211-266 / reduced_kkt.c:31-51 */
static void pcg_update_precond(oracle_linsys* s) {
  int j, k;
  if (s->precond_type == OSQP_NO_PRECONDITIONER) {
    for (j = 0; j < s->n; j++) s->precond_inv[j] = 1.0;
    return;
  }
  for (j = 0; j < s->n; j++) {
    double d = 0.0, at = 0.0;
    for (k = s->P->p[j]; k < s->P->p[j + 1]; k++) if (s->P->i[k] == j) d = s->P->x[k];
    for (k = s->A->p[j]; k < s->A->p[j + 1]; k++) {
      double a = s->A->x[k];
      at += a * a * (s->rho_vec ? s->rho_vec[s->A->i[k]] : s->rho_sc);
    }
    s->precond_inv[j] = 1.0 / (s->sigma + d + at);
  }
}

/* cuda_pcg_interface.cu:32-64 */
static double pcg_tolerance(oracle_linsys* s, int admm_iter, double rhs_norm) {
  double eps;
  if (admm_iter == 1) {
    s->reduction_factor = s->tol_fraction;
    if (rhs_norm < OSQP_CG_TOL_MIN) s->eps_prev = 1.0;
    else s->eps_prev = rhs_norm * s->reduction_factor;
    return s->eps_prev;
  }
  if (s->zero_pcg_iters >= s->reduction_threshold) {
    s->reduction_factor /= 2;
    s->zero_pcg_iters = 0;
  }
  eps = s->reduction_factor * sqrt((*s->scaled_prim_res) * (*s->scaled_dual_res));
  eps = c_max(c_min(eps, s->eps_prev), OSQP_CG_TOL_MIN);
  s->eps_prev = eps;
  return eps;
}

/* cuda_pcg.cu:113-208 (x is warm-started from the previous solve) */
static int pcg_run(oracle_linsys* s, double eps) {
  int n = s->n, i, iter = 0;
  double rTy, rTy_prev, pKp, alpha, beta, rnorm;
  pcg_matvec(s, s->px, s->pr);
  for (i = 0; i < n; i++) s->pr[i] -= s->rhs[i];
  rnorm = vnorm_inf(s->pr, n);
  for (i = 0; i < n; i++) s->py[i] = s->precond_inv[i] * s->pr[i];
  for (i = 0; i < n; i++) s->pp[i] = -s->py[i];
  rTy = vdot(s->py, s->pr, n);
  while (rnorm > eps && iter < s->max_iter) {
    pcg_matvec(s, s->pp, s->pKp);
    pKp = vdot(s->pp, s->pKp, n);
    alpha = rTy / pKp;
    for (i = 0; i < n; i++) s->px[i] += alpha * s->pp[i];
    for (i = 0; i < n; i++) s->pr[i] += alpha * s->pKp[i];
    for (i = 0; i < n; i++) s->py[i] = s->precond_inv[i] * s->pr[i];
    rTy_prev = rTy;
    rTy = vdot(s->py, s->pr, n);
    rnorm = vnorm_inf(s->pr, n);
    beta = rTy / rTy_prev;
    for (i = 0; i < n; i++) s->pp[i] = beta * s->pp[i] - s->py[i];
    iter++;
  }
  return iter;
}

/* cuda_pcg_interface.cu:67-92, 229-273 */
static void indirect_solve(oracle_linsys* s, double* b, int admm_iter) {
  int n = s->n, m = s->m, i, iters;
  double eps;
  for (i = 0; i < n; i++) s->rhs[i] = b[i];
  if (m) {
    for (i = 0; i < m; i++) s->pz[i] = b[n + i] * (s->rho_vec ? s->rho_vec[i] : s->rho_sc);
    csc_Atxpy(s->A, s->pz, s->rhs, 1.0, 1.0);
  }
  eps = pcg_tolerance(s, admm_iter, vnorm_inf(s->rhs, n));
  iters = pcg_run(s, eps);
  s->total_pcg_iters += iters;
  for (i = 0; i < n; i++) b[i] = s->px[i];
  if (m) {
    csc_Axpy(s->A, s->px, s->pz, 1.0, 0.0);
    for (i = 0; i < m; i++) b[n + i] = s->pz[i];
  }
  if (iters == 0) s->zero_pcg_iters++;
  else s->zero_pcg_iters = 0;
}

static int linsys_init(oracle_linsys** sp, const ocsc* P, const ocsc* A, const double* rho_vec,
                       const OSQPSettings* st, const double* spr, const double* sdr) {
  oracle_linsys* s = (oracle_linsys*)calloc(1, sizeof(oracle_linsys));
  *sp = s;
  s->type = st->linsys_solver;
  s->n = P->n;
  s->m = A->m;
  s->P = P;
  s->A = A;
  s->rho_vec = rho_vec;
  s->sigma = st->sigma;
  s->rho_sc = st->rho;
  if (s->type == OSQP_INDIRECT_SOLVER) {
    int n = s->n, m = s->m;
    s->max_iter = st->cg_max_iter;
    s->precond_type = st->cg_precond;
    s->reduction_threshold = st->cg_tol_reduction;
    s->tol_fraction = st->cg_tol_fraction;
    s->reduction_factor = st->cg_tol_fraction;
    s->scaled_prim_res = spr;
    s->scaled_dual_res = sdr;
    s->px = (double*)calloc((size_t)n + 1, sizeof(double));
    s->pp = (double*)calloc((size_t)n + 1, sizeof(double));
    s->pKp = (double*)calloc((size_t)n + 1, sizeof(double));
    s->py = (double*)calloc((size_t)n + 1, sizeof(double));
    s->pr = (double*)calloc((size_t)n + 1, sizeof(double));
    s->rhs = (double*)calloc((size_t)n + 1, sizeof(double));
    s->pz = (double*)calloc((size_t)m + 1, sizeof(double));
    s->precond_inv = (double*)calloc((size_t)n + 1, sizeof(double));
    pcg_update_precond(s);
    return 0;
  }
  return kkt_assemble_factor(s, 1);
}

static void linsys_free(oracle_linsys* s) {
  if (!s) return;
  free(s->Kp); free(s->Ki); free(s->Kx); free(s->perm); free(s->pinv);
  free(s->rho_pos); free(s->sigma_pos); free(s->bp); free(s->sol);
  ldl_free(s->F);
  free(s->px); free(s->pp); free(s->pKp); free(s->py); free(s->pr); free(s->rhs); free(s->pz);
  free(s->precond_inv);
  free(s);
}

static void linsys_solve(oracle_linsys* s, double* b, int admm_iter) {
  if (s->type == OSQP_INDIRECT_SOLVER) indirect_solve(s, b, admm_iter);
  else direct_solve(s, b);
}

static void linsys_warm_start(oracle_linsys* s, const double* x) {
  if (s->type == OSQP_INDIRECT_SOLVER) memcpy(s->px, x, sizeof(double) * (size_t)s->n);
}

static int linsys_update_rho(oracle_linsys* s, double rho_sc) {
  s->rho_sc = rho_sc;
  if (s->type == OSQP_INDIRECT_SOLVER) { pcg_update_precond(s); return 0; }
  {
    int i;
    for (i = 0; i < s->m; i++)
      s->Kx[s->rho_pos[i]] = s->rho_vec ? -1.0 / s->rho_vec[i] : -1.0 / s->rho_sc;
    return kkt_assemble_factor(s, 0);
  }
}

static int linsys_update_matrices(oracle_linsys* s) {
  if (s->type == OSQP_INDIRECT_SOLVER) { pcg_update_precond(s); return 0; }
  return kkt_refresh(s);
}

static void linsys_update_settings(oracle_linsys* s, const OSQPSettings* st) {
  if (s->type != OSQP_INDIRECT_SOLVER) return;
  s->max_iter = st->cg_max_iter;
  s->reduction_threshold = st->cg_tol_reduction;
  s->tol_fraction = st->cg_tol_fraction;
  if (s->precond_type != (int)st->cg_precond) {
    s->precond_type = st->cg_precond;
    pcg_update_precond(s);
  }
}

/* ========================================================================
 * Solver (src/osqp_api.c, src/auxil.c, src/scaling.c)
 * ====================================================================== */

typedef struct {
  OSQPSettings settings;
  OSQPInfo info;
  OSQPSolution solution;
  int has_solution_buf;
  int n, m;
  ocsc *P, *A;
  double *q, *l, *u;
  /* scaling */
  int scaled;
  double c, cinv;
  double *D, *Dinv, *E, *Einv, *D_temp, *D_temp_A, *E_temp;
  /* rho */
  double *rho_vec, *rho_inv_vec, rho_inv;
  int* constr_type;
  /* iterates */
  double *x, *y, *z, *xz_tilde, *x_prev, *z_prev;
  double *Ax, *Px, *Aty, *delta_y, *Atdelta_y, *delta_x, *Pdelta_x, *Adelta_x;
  double xtPx, qtx, SC, scaled_dual_gap, scaled_prim_res, scaled_dual_res, last_rel_kkt;
  int rho_updated;
  oracle_linsys* linsys;
  /* timer (OSQP_ENABLE_PROFILING) */
  struct timespec tic;
  int first_run, clear_update_time, rho_update_from_solve;
} OracleSolver;

static void tic(OracleSolver* s) { clock_gettime(CLOCK_MONOTONIC, &s->tic); }
static double toc(OracleSolver* s) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)(t.tv_sec - s->tic.tv_sec) + 1e-9 * (double)(t.tv_nsec - s->tic.tv_nsec);
}

static double* vnew(int n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

static const char* STATUS_MSG[] = {"", "solved", "solved inaccurate", "primal infeasible",
                                   "primal infeasible inaccurate", "dual infeasible",
                                   "dual infeasible inaccurate", "maximum iterations reached",
                                   "run time limit reached", "problem non convex", "interrupted",
                                   "unsolved"};

static void update_status(OSQPInfo* info, int v) {
  info->status_val = v;
  strcpy(info->status, STATUS_MSG[v]);
}

/* auxil.c:76-105 + vector.c:885-921 (ew_bounds_type) */
static int set_rho_vec(OracleSolver* s) {
  int i, changed = 0;
  double tol = OSQP_RHO_TOL, infval = OSQP_INFTY * OSQP_MIN_SCALING;
  s->settings.rho = c_min(c_max(s->settings.rho, OSQP_RHO_MIN), OSQP_RHO_MAX);
  for (i = 0; i < s->m; i++) {
    int old = s->constr_type[i], t;
    if ((s->l[i] < -infval) && (s->u[i] > infval)) t = -1;
    else if (s->u[i] - s->l[i] < tol) t = 1;
    else t = 0;
    s->constr_type[i] = t;
    changed = changed || (t != old);
  }
  for (i = 0; i < s->m; i++) {
    if (s->constr_type[i] == 0) s->rho_vec[i] = s->settings.rho;
    else if (s->constr_type[i] > 0) s->rho_vec[i] = OSQP_RHO_EQ_OVER_RHO_INEQ * s->settings.rho;
    else s->rho_vec[i] = OSQP_RHO_MIN;
  }
  for (i = 0; i < s->m; i++) s->rho_inv_vec[i] = 1.0 / s->rho_vec[i];
  return changed;
}

/* scaling.c:8-13 */
static double limit_scaling_scalar(double v) {
  v = v < OSQP_MIN_SCALING ? 1.0 : v;
  v = v > OSQP_MAX_SCALING ? OSQP_MAX_SCALING : v;
  return v;
}

static void limit_scaling_vector(double* v, int n) {
  int i;
  for (i = 0; i < n; i++) v[i] = v[i] < OSQP_MIN_SCALING ? 1.0 : v[i];
  for (i = 0; i < n; i++) v[i] = v[i] > OSQP_MAX_SCALING ? OSQP_MAX_SCALING : v[i];
}

/* scaling.c:48-164 */
static void scale_data(OracleSolver* s) {
  int i, k, n = s->n, m = s->m;
  double c_temp, inf_norm_q;
  s->c = 1.0;
  for (i = 0; i < n; i++) { s->D[i] = 1.0; s->Dinv[i] = 1.0; }
  for (i = 0; i < m; i++) { s->E[i] = 1.0; s->Einv[i] = 1.0; }
  for (k = 0; k < s->settings.scaling; k++) {
    csc_col_norm_inf(s->P, s->D_temp);
    csc_col_norm_inf(s->A, s->D_temp_A);
    for (i = 0; i < n; i++) s->D_temp[i] = c_max(s->D_temp[i], s->D_temp_A[i]);
    csc_row_norm_inf(s->A, s->E_temp);
    limit_scaling_vector(s->D_temp, n);
    limit_scaling_vector(s->E_temp, m);
    for (i = 0; i < n; i++) s->D_temp[i] = sqrt(s->D_temp[i]);
    for (i = 0; i < m; i++) s->E_temp[i] = sqrt(s->E_temp[i]);
    for (i = 0; i < n; i++) s->D_temp[i] = 1.0 / s->D_temp[i];
    for (i = 0; i < m; i++) s->E_temp[i] = 1.0 / s->E_temp[i];
    csc_lmult_diag(s->P, s->D_temp);
    csc_rmult_diag(s->P, s->D_temp);
    csc_lmult_diag(s->A, s->E_temp);
    csc_rmult_diag(s->A, s->D_temp);
    for (i = 0; i < n; i++) s->q[i] *= s->D_temp[i];
    for (i = 0; i < n; i++) s->D[i] *= s->D_temp[i];
    for (i = 0; i < m; i++) s->E[i] *= s->E_temp[i];
    csc_col_norm_inf(s->P, s->D_temp);
    c_temp = 0.0;
    for (i = 0; i < n; i++) c_temp += c_absval(s->D_temp[i]);
    c_temp = c_temp / n;
    inf_norm_q = vnorm_inf(s->q, n);
    inf_norm_q = limit_scaling_scalar(inf_norm_q);
    c_temp = c_max(c_temp, inf_norm_q);
    c_temp = limit_scaling_scalar(c_temp);
    c_temp = 1. / c_temp;
    csc_scale(s->P, c_temp);
    for (i = 0; i < n; i++) s->q[i] *= c_temp;
    s->c *= c_temp;
  }
  s->cinv = 1. / s->c;
  for (i = 0; i < n; i++) s->Dinv[i] = 1.0 / s->D[i];
  for (i = 0; i < m; i++) s->Einv[i] = 1.0 / s->E[i];
  for (i = 0; i < m; i++) s->l[i] *= s->E[i];
  for (i = 0; i < m; i++) s->u[i] *= s->E[i];
}

/* scaling.c:169-192 */
static void unscale_data(OracleSolver* s) {
  int i;
  csc_scale(s->P, s->cinv);
  csc_lmult_diag(s->P, s->Dinv);
  csc_rmult_diag(s->P, s->Dinv);
  for (i = 0; i < s->n; i++) s->q[i] *= s->cinv;
  for (i = 0; i < s->n; i++) s->q[i] *= s->Dinv[i];
  csc_lmult_diag(s->A, s->Einv);
  csc_rmult_diag(s->A, s->Dinv);
  for (i = 0; i < s->m; i++) s->l[i] *= s->Einv[i];
  for (i = 0; i < s->m; i++) s->u[i] *= s->Einv[i];
}

static void cold_start(OracleSolver* s) {
  memset(s->x, 0, sizeof(double) * (size_t)s->n);
  memset(s->z, 0, sizeof(double) * (size_t)(s->m));
  memset(s->y, 0, sizeof(double) * (size_t)(s->m));
  linsys_warm_start(s->linsys, s->x);
}

/* validation: auxil.c:950-1017, 1040-1223 */
static int validate_data(const OSQPCscMatrix* P, const double* q, const OSQPCscMatrix* A,
                         const double* l, const double* u, int m, int n) {
  int j, ptr;
  if (!P || !A || !q) return 1;
  if ((n <= 0) || (m < 0)) return 1;
  if (P->m != n || P->m != P->n) return 1;
  for (j = 0; j < n; j++)
    for (ptr = P->p[j]; ptr < P->p[j + 1]; ptr++)
      if (P->i[ptr] > j) return 1;
  if ((A->m != m) || (A->n != n)) return 1;
  for (j = 0; j < m; j++) if (l[j] > u[j]) return 1;
  return 0;
}

static int validate_settings(const OSQPSettings* st, int from_setup) {
  if (!st) return 1;
  if (from_setup && st->linsys_solver != OSQP_DIRECT_SOLVER && st->linsys_solver != OSQP_INDIRECT_SOLVER) return 1;
  if (from_setup && st->allocate_solution != 0 && st->allocate_solution != 1) return 1;
  if (st->verbose != 0 && st->verbose != 1) return 1;
  if (st->profiler_level < 0 || st->profiler_level > 2) return 1;
  if (st->warm_starting != 0 && st->warm_starting != 1) return 1;
  if (from_setup && st->scaling < 0) return 1;
  if (st->polishing != 0 && st->polishing != 1) return 1;
  if (from_setup && st->rho <= 0.0) return 1;
  if (from_setup && st->rho_is_vec != 0 && st->rho_is_vec != 1) return 1;
  if (from_setup && st->sigma <= 0.0) return 1;
  if (st->alpha <= 0.0 || st->alpha >= 2.0) return 1;
  if (st->cg_max_iter <= 0) return 1;
  if (st->cg_tol_reduction <= 0) return 1;
  if (st->cg_tol_fraction <= 0.0 || st->cg_tol_fraction >= 1.0) return 1;
  if ((from_setup && st->adaptive_rho < 0) || st->adaptive_rho >= _OSQP_ADAPTIVE_RHO_UPDATE_LAST_VALUE) return 1;
  if (from_setup && st->adaptive_rho_interval < 0) return 1;
  if (from_setup && st->adaptive_rho_fraction <= 0) return 1;
  if (from_setup && st->adaptive_rho_tolerance < 1.0) return 1;
  if (st->max_iter <= 0) return 1;
  if (st->eps_abs < 0.0 || st->eps_rel < 0.0) return 1;
  if (st->eps_rel == 0.0 && st->eps_abs == 0.0) return 1;
  if (st->eps_prim_inf <= 0.0 || st->eps_dual_inf <= 0.0) return 1;
  if (st->scaled_termination != 0 && st->scaled_termination != 1) return 1;
  if (st->check_termination < 0) return 1;
  if (st->check_dualgap != 0 && st->check_dualgap != 1) return 1;
  if (st->time_limit <= 0.0) return 1;
  if (st->delta <= 0.0) return 1;
  if (st->polish_refine_iter < 0) return 1;
  return 0;
}

void oracle_set_default_settings(OSQPSettings* st) {
  st->device = 0;
  st->linsys_solver = OSQP_DIRECT_SOLVER;
  st->allocate_solution = 1;
  st->profiler_level = 0;
  st->verbose = 0;
  st->warm_starting = OSQP_WARM_STARTING;
  st->scaling = OSQP_SCALING;
  st->polishing = OSQP_POLISHING;
  st->rho = OSQP_RHO;
  st->rho_is_vec = OSQP_RHO_IS_VEC;
  st->sigma = OSQP_SIGMA;
  st->alpha = OSQP_ALPHA;
  st->cg_max_iter = OSQP_CG_MAX_ITER;
  st->cg_tol_reduction = OSQP_CG_TOL_REDUCTION;
  st->cg_tol_fraction = OSQP_CG_TOL_FRACTION;
  st->cg_precond = OSQP_DIAGONAL_PRECONDITIONER;
  st->adaptive_rho = OSQP_ADAPTIVE_RHO_UPDATE_DEFAULT;
  st->adaptive_rho_interval = OSQP_ADAPTIVE_RHO_INTERVAL;
  st->adaptive_rho_fraction = OSQP_ADAPTIVE_RHO_FRACTION;
  st->adaptive_rho_tolerance = OSQP_ADAPTIVE_RHO_TOLERANCE;
  st->max_iter = OSQP_MAX_ITER;
  st->eps_abs = OSQP_EPS_ABS;
  st->eps_rel = OSQP_EPS_REL;
  st->eps_prim_inf = OSQP_EPS_PRIM_INF;
  st->eps_dual_inf = OSQP_EPS_DUAL_INF;
  st->scaled_termination = OSQP_SCALED_TERMINATION;
  st->check_termination = OSQP_CHECK_TERMINATION;
  st->check_dualgap = OSQP_CHECK_DUALGAP;
  st->time_limit = OSQP_TIME_LIMIT;
  st->delta = OSQP_DELTA;
  st->polish_refine_iter = OSQP_POLISH_REFINE_ITER;
}

void oracle_cleanup(OracleSolver* s);

/* osqp_api.c:332-631 */
int oracle_setup(OracleSolver** sp, const OSQPCscMatrix* P, const double* q,
                 const OSQPCscMatrix* A, const double* l, const double* u,
                 int m, int n, const OSQPSettings* st) {
  OracleSolver* s;
  int flag;
  *sp = NULL;
  if (validate_data(P, q, A, l, u, m, n)) return OSQP_DATA_VALIDATION_ERROR;
  if (validate_settings(st, 1)) return OSQP_SETTINGS_VALIDATION_ERROR;
  s = (OracleSolver*)calloc(1, sizeof(OracleSolver));
  *sp = s;
  tic(s);
  s->n = n;
  s->m = m;
  s->settings = *st;
  s->P = ocsc_copy(P);
  s->A = ocsc_copy(A);
  s->q = vnew(n); memcpy(s->q, q, sizeof(double) * (size_t)n);
  s->l = vnew(m); if (m) memcpy(s->l, l, sizeof(double) * (size_t)m);
  s->u = vnew(m); if (m) memcpy(s->u, u, sizeof(double) * (size_t)m);
  if (st->rho_is_vec) {
    s->rho_vec = vnew(m);
    s->rho_inv_vec = vnew(m);
    s->constr_type = (int*)calloc((size_t)(m > 0 ? m : 1), sizeof(int));
  }
  s->x = vnew(n); s->z = vnew(m); s->xz_tilde = vnew(n + m);
  s->x_prev = vnew(n); s->z_prev = vnew(m); s->y = vnew(m);
  s->Ax = vnew(m); s->Px = vnew(n); s->Aty = vnew(n);
  s->delta_y = vnew(m); s->Atdelta_y = vnew(n);
  s->delta_x = vnew(n); s->Pdelta_x = vnew(n); s->Adelta_x = vnew(m);
  s->D = vnew(n); s->Dinv = vnew(n); s->E = vnew(m); s->Einv = vnew(m);
  s->D_temp = vnew(n); s->D_temp_A = vnew(n); s->E_temp = vnew(m);
  s->scaled = st->scaling ? 1 : 0;
  if (st->scaling) scale_data(s);
  if (st->rho_is_vec) {
    set_rho_vec(s);
  } else {
    s->settings.rho = c_min(c_max(st->rho, OSQP_RHO_MIN), OSQP_RHO_MAX);
    s->rho_inv = 1. / st->rho;
  }
  flag = linsys_init(&s->linsys, s->P, s->A, s->rho_vec, &s->settings,
                     &s->scaled_prim_res, &s->scaled_dual_res);
  if (flag == OSQP_NONCVX_ERROR) {
    update_status(&s->info, OSQP_NON_CVX);
    return flag;
  } else if (flag) {
    return flag;
  }
  cold_start(s);
  if (st->allocate_solution) {
    s->has_solution_buf = 1;
    s->solution.x = vnew(n);
    s->solution.y = vnew(m);
    s->solution.prim_inf_cert = vnew(m);
    s->solution.dual_inf_cert = vnew(n);
  }
  s->info.status_polish = OSQP_POLISH_NOT_PERFORMED;
  update_status(&s->info, OSQP_UNSOLVED);
  s->info.solve_time = 0.0;
  s->info.update_time = 0.0;
  s->info.polish_time = 0.0;
  s->info.run_time = 0.0;
  s->info.setup_time = toc(s);
  s->first_run = 1;
  s->clear_update_time = 0;
  s->rho_update_from_solve = 0;
  s->info.rho_updates = 0;
  s->info.rho_estimate = s->settings.rho;
  s->info.obj_val = OSQP_INFTY;
  s->info.prim_res = OSQP_INFTY;
  s->info.dual_res = OSQP_INFTY;
  s->info.rel_kkt_error = OSQP_INFTY;
  s->last_rel_kkt = OSQP_INFTY;
  s->rho_updated = 0;
  if (s->settings.adaptive_rho == OSQP_ADAPTIVE_RHO_UPDATE_ITERATIONS && s->settings.adaptive_rho_interval == 0) {
    s->settings.adaptive_rho_interval = s->settings.check_termination
        ? OSQP_ADAPTIVE_RHO_MULTIPLE_TERMINATION * s->settings.check_termination
        : OSQP_ADAPTIVE_RHO_FIXED;
  }
  if (s->settings.adaptive_rho == OSQP_ADAPTIVE_RHO_UPDATE_KKT_ERROR && s->settings.adaptive_rho_interval == 0)
    s->settings.adaptive_rho_interval = 1;
  return 0;
}

/* auxil.c:136-229 */
static void compute_rhs(OracleSolver* s) {
  int i, n = s->n, m = s->m;
  double* xt = s->xz_tilde;
  double* zt = s->xz_tilde + n;
  for (i = 0; i < n; i++) xt[i] = s->settings.sigma * s->x_prev[i] + -1. * s->q[i];
  if (s->settings.rho_is_vec) {
    for (i = 0; i < m; i++) zt[i] = s->rho_inv_vec[i] * s->y[i];
    for (i = 0; i < m; i++) zt[i] = -1.0 * zt[i] + 1.0 * s->z_prev[i];
  } else {
    for (i = 0; i < m; i++) zt[i] = 1.0 * s->z_prev[i] + -s->rho_inv * s->y[i];
  }
}

static void update_x(OracleSolver* s) {
  int i;
  double a = s->settings.alpha;
  for (i = 0; i < s->n; i++) s->x[i] = a * s->xz_tilde[i] + (1.0 - a) * s->x_prev[i];
  for (i = 0; i < s->n; i++) s->delta_x[i] = s->x[i] - s->x_prev[i];
}

static void update_z(OracleSolver* s) {
  int i;
  double a = s->settings.alpha;
  const double* zt = s->xz_tilde + s->n;
  if (s->settings.rho_is_vec) {
    for (i = 0; i < s->m; i++) s->z[i] = s->rho_inv_vec[i] * s->y[i];
    for (i = 0; i < s->m; i++) s->z[i] += a * zt[i] + (1.0 - a) * s->z_prev[i];
  } else {
    for (i = 0; i < s->m; i++) s->z[i] = a * zt[i] + (1.0 - a) * s->z_prev[i] + s->rho_inv * s->y[i];
  }
  for (i = 0; i < s->m; i++) s->z[i] = c_min(c_max(s->z[i], s->l[i]), s->u[i]);
}

static void update_y(OracleSolver* s) {
  int i;
  double a = s->settings.alpha;
  const double* zt = s->xz_tilde + s->n;
  for (i = 0; i < s->m; i++) s->delta_y[i] = a * zt[i] + (1.0 - a) * s->z_prev[i] + -1.0 * s->z[i];
  if (s->settings.rho_is_vec) for (i = 0; i < s->m; i++) s->delta_y[i] *= s->rho_vec[i];
  else for (i = 0; i < s->m; i++) s->delta_y[i] *= s->settings.rho;
  for (i = 0; i < s->m; i++) s->y[i] += s->delta_y[i];
}

/* vector.c:676-702 */
static void project_polar_reccone(double* y, const double* l, const double* u, int n, double infval) {
  int i;
  for (i = 0; i < n; i++) {
    if (u[i] > +infval) {
      if (l[i] < -infval) y[i] = 0.0;
      else y[i] = c_min(y[i], 0.0);
    } else if (l[i] < -infval) {
      y[i] = c_max(y[i], 0.0);
    }
  }
}

static double dot_signed(const double* a, const double* b, int n, int sign) {
  int i;
  double d = 0.0;
  if (sign == 1) for (i = 0; i < n; i++) d += a[i] * c_max(b[i], 0.);
  else if (sign == -1) for (i = 0; i < n; i++) d += a[i] * c_min(b[i], 0.);
  else d = vdot(a, b, n);
  return d;
}

static double scaled_norm_inf(const double* S, const double* v, int n) {
  int i;
  double r = 0.0;
  for (i = 0; i < n; i++) { double a = c_absval(S[i] * v[i]); if (a > r) r = a; }
  return r;
}

/* auxil.c:231-285 */
static void compute_obj_val_dual_gap(OracleSolver* s, const double* x, const double* y,
                                     double* pobj, double* dobj, double* gap) {
  int i;
  double quad = vdot(s->Px, x, s->n);
  double lin = vdot(s->q, x, s->n);
  double sup;
  memcpy(s->z_prev, y, sizeof(double) * (size_t)s->m);
  project_polar_reccone(s->z_prev, s->l, s->u, s->m, OSQP_INFTY * OSQP_MIN_SCALING);
  for (i = 0; i < s->m; i++) if (c_absval(s->z_prev[i]) < OSQP_ZERO_DEADZONE) s->z_prev[i] = 0.0;
  sup = dot_signed(s->u, s->z_prev, s->m, +1);
  sup += dot_signed(s->l, s->z_prev, s->m, -1);
  *pobj = 0.5 * quad + lin;
  *dobj = -0.5 * quad - sup;
  s->scaled_dual_gap = quad + lin + sup;
  if (s->settings.scaling) {
    *pobj *= s->cinv;
    *dobj *= s->cinv;
    *gap = s->cinv * s->scaled_dual_gap;
  } else {
    *gap = s->scaled_dual_gap;
  }
  s->xtPx = quad;
  s->qtx = lin;
  s->SC = sup;
}

static int use_unscaled(const OracleSolver* s) { return s->settings.scaling && !s->settings.scaled_termination; }

/* auxil.c:308-332 */
static double compute_prim_res(OracleSolver* s, const double* x, const double* z) {
  int i;
  csc_Axpy(s->A, x, s->Ax, 1.0, 0.0);
  for (i = 0; i < s->m; i++) s->z_prev[i] = s->Ax[i] - z[i];
  s->scaled_prim_res = vnorm_inf(s->z_prev, s->m);
  if (use_unscaled(s)) return scaled_norm_inf(s->Einv, s->z_prev, s->m);
  return s->scaled_prim_res;
}

/* auxil.c:371-410 */
static double compute_dual_res(OracleSolver* s, const double* x, const double* y) {
  int i;
  memcpy(s->x_prev, s->q, sizeof(double) * (size_t)s->n);
  csc_Axpy_sym_triu(s->P, x, s->Px, 1.0, 0.0);
  for (i = 0; i < s->n; i++) s->x_prev[i] += s->Px[i];
  if (s->m) {
    csc_Atxpy(s->A, y, s->Aty, 1.0, 0.0);
    for (i = 0; i < s->n; i++) s->x_prev[i] += s->Aty[i];
  }
  s->scaled_dual_res = vnorm_inf(s->x_prev, s->n);
  if (use_unscaled(s)) return s->cinv * scaled_norm_inf(s->Dinv, s->x_prev, s->n);
  return s->scaled_dual_res;
}

static double compute_prim_tol(OracleSolver* s, double eps_abs, double eps_rel) {
  double a, b;
  if (use_unscaled(s)) {
    a = scaled_norm_inf(s->Einv, s->z, s->m);
    b = scaled_norm_inf(s->Einv, s->Ax, s->m);
  } else {
    a = vnorm_inf(s->z, s->m);
    b = vnorm_inf(s->Ax, s->m);
  }
  return eps_abs + eps_rel * c_max(a, b);
}

static double compute_dual_tol(OracleSolver* s, double eps_abs, double eps_rel) {
  double r;
  if (use_unscaled(s)) {
    r = scaled_norm_inf(s->Dinv, s->q, s->n);
    r = c_max(r, scaled_norm_inf(s->Dinv, s->Aty, s->n));
    r = c_max(r, scaled_norm_inf(s->Dinv, s->Px, s->n));
    r *= s->cinv;
  } else {
    r = vnorm_inf(s->q, s->n);
    r = c_max(r, vnorm_inf(s->Aty, s->n));
    r = c_max(r, vnorm_inf(s->Px, s->n));
  }
  return eps_abs + eps_rel * r;
}

static double compute_gap_tol(OracleSolver* s, double eps_abs, double eps_rel) {
  double r = c_absval(s->xtPx);
  r = c_max(r, c_absval(s->qtx));
  r = c_max(r, c_absval(s->SC));
  if (use_unscaled(s)) r = s->cinv * r;
  return eps_abs + eps_rel * r;
}

/* auxil.c:460-512 */
static int is_primal_infeasible(OracleSolver* s, double eps) {
  int i;
  double norm_dy, lhs;
  project_polar_reccone(s->delta_y, s->l, s->u, s->m, OSQP_INFTY * OSQP_MIN_SCALING);
  if (use_unscaled(s)) {
    for (i = 0; i < s->m; i++) s->Adelta_x[i] = s->E[i] * s->delta_y[i];
    norm_dy = vnorm_inf(s->Adelta_x, s->m);
  } else {
    norm_dy = vnorm_inf(s->delta_y, s->m);
  }
  if (norm_dy > OSQP_DIVISION_TOL) {
    lhs = dot_signed(s->u, s->delta_y, s->m, +1);
    lhs += dot_signed(s->l, s->delta_y, s->m, -1);
    if (lhs < 0.0) {
      csc_Atxpy(s->A, s->delta_y, s->Atdelta_y, 1.0, 0.0);
      if (use_unscaled(s)) for (i = 0; i < s->n; i++) s->Atdelta_y[i] *= s->Dinv[i];
      return vnorm_inf(s->Atdelta_y, s->n) < eps * norm_dy;
    }
  }
  return 0;
}

/* auxil.c:514-588 */
static int is_dual_infeasible(OracleSolver* s, double eps) {
  int i;
  double norm_dx, cost_scaling;
  if (use_unscaled(s)) {
    norm_dx = scaled_norm_inf(s->D, s->delta_x, s->n);
    cost_scaling = s->c;
  } else {
    norm_dx = vnorm_inf(s->delta_x, s->n);
    cost_scaling = 1.0;
  }
  if (norm_dx > OSQP_DIVISION_TOL) {
    if (vdot(s->q, s->delta_x, s->n) < 0.0) {
      csc_Axpy_sym_triu(s->P, s->delta_x, s->Pdelta_x, 1.0, 0.0);
      if (use_unscaled(s)) for (i = 0; i < s->n; i++) s->Pdelta_x[i] *= s->Dinv[i];
      if (vnorm_inf(s->Pdelta_x, s->n) < cost_scaling * eps * norm_dx) {
        double infval = OSQP_INFTY * OSQP_MIN_SCALING, tol = eps * norm_dx;
        csc_Axpy(s->A, s->delta_x, s->Adelta_x, 1.0, 0.0);
        if (use_unscaled(s)) for (i = 0; i < s->m; i++) s->Adelta_x[i] *= s->Einv[i];
        for (i = 0; i < s->m; i++) {
          if (((s->u[i] < +infval) && (s->Adelta_x[i] > +tol)) ||
              ((s->l[i] > -infval) && (s->Adelta_x[i] < -tol)))
            return 0;
        }
        return 1;
      }
    }
  }
  return 0;
}

static int has_solution(const OSQPInfo* info) {
  return info->status_val != OSQP_PRIMAL_INFEASIBLE && info->status_val != OSQP_PRIMAL_INFEASIBLE_INACCURATE &&
         info->status_val != OSQP_DUAL_INFEASIBLE && info->status_val != OSQP_DUAL_INFEASIBLE_INACCURATE &&
         info->status_val != OSQP_NON_CVX;
}

/* auxil.c:676-762 */
static void update_info(OracleSolver* s, int iter) {
  OSQPInfo* info = &s->info;
  info->iter = iter;
  if (s->m == 0) info->prim_res = 0.;
  else info->prim_res = compute_prim_res(s, s->x, s->z);
  info->dual_res = compute_dual_res(s, s->x, s->y);
  compute_obj_val_dual_gap(s, s->x, s->y, &info->obj_val, &info->dual_obj_val, &info->duality_gap);
  info->primdual_int += c_absval(info->duality_gap);
  info->solve_time = toc(s);
  info->rel_kkt_error = c_max(c_max(info->dual_res, info->prim_res), info->duality_gap);
}

static void reset_info(OSQPInfo* info) {
  info->solve_time = 0.0;
  info->polish_time = 0.0;
  update_status(info, OSQP_UNSOLVED);
  info->rho_updates = 0;
}

/* auxil.c:808-945 */
static int check_termination(OracleSolver* s, int approximate) {
  OSQPInfo* info = &s->info;
  double eps_abs = s->settings.eps_abs, eps_rel = s->settings.eps_rel;
  double eps_pinf = s->settings.eps_prim_inf, eps_dinf = s->settings.eps_dual_inf;
  int prc = 0, drc = 0, pic = 0, dic = 0, gapc = 0, i;
  if ((info->prim_res > OSQP_INFTY) || (info->dual_res > OSQP_INFTY)) {
    update_status(info, OSQP_NON_CVX);
    info->obj_val = OSQP_NAN;
    return 1;
  }
  if (approximate) { eps_abs *= 10; eps_rel *= 10; eps_pinf *= 10; eps_dinf *= 10; }
  if (s->m == 0) {
    prc = 1;
  } else {
    if (info->prim_res < compute_prim_tol(s, eps_abs, eps_rel)) prc = 1;
    else pic = is_primal_infeasible(s, eps_pinf);
  }
  if (info->dual_res < compute_dual_tol(s, eps_abs, eps_rel)) drc = 1;
  else dic = is_dual_infeasible(s, eps_dinf);
  if (s->settings.check_dualgap) {
    double eg = compute_gap_tol(s, eps_abs, eps_rel);
    if (use_unscaled(s)) { if (c_absval(info->duality_gap) < eg) gapc = 1; }
    else { if (c_absval(s->scaled_dual_gap) < eg) gapc = 1; }
  } else {
    gapc = 1;
  }
  if (prc && drc && gapc) {
    update_status(info, approximate ? OSQP_SOLVED_INACCURATE : OSQP_SOLVED);
    return 1;
  } else if (pic) {
    update_status(info, approximate ? OSQP_PRIMAL_INFEASIBLE_INACCURATE : OSQP_PRIMAL_INFEASIBLE);
    if (use_unscaled(s)) for (i = 0; i < s->m; i++) s->delta_y[i] *= s->E[i];
    info->obj_val = OSQP_INFTY;
    return 1;
  } else if (dic) {
    update_status(info, approximate ? OSQP_DUAL_INFEASIBLE_INACCURATE : OSQP_DUAL_INFEASIBLE);
    if (use_unscaled(s)) for (i = 0; i < s->n; i++) s->delta_x[i] *= s->D[i];
    info->obj_val = -OSQP_INFTY;
    return 1;
  }
  return 0;
}

/* auxil.c:14-47 */
static double compute_rho_estimate(OracleSolver* s) {
  double pr = s->scaled_prim_res, dr = s->scaled_dual_res, pn, dn, est;
  pn = c_max(vnorm_inf(s->z, s->m), vnorm_inf(s->Ax, s->m));
  pr /= (pn + OSQP_DIVISION_TOL);
  dn = vnorm_inf(s->q, s->n);
  dn = c_max(dn, vnorm_inf(s->Aty, s->n));
  dn = c_max(dn, vnorm_inf(s->Px, s->n));
  dr /= (dn + OSQP_DIVISION_TOL);
  est = s->settings.rho * sqrt(pr / dr);
  return c_min(c_max(est, OSQP_RHO_MIN), OSQP_RHO_MAX);
}

/* osqp_api.c:1412-1464 */
int oracle_update_rho(OracleSolver* s, double rho_new) {
  int i;
  double t0 = 0.0;
  if (rho_new <= 0) return 1;
  if (!s->rho_update_from_solve) {
    if (s->clear_update_time) { s->clear_update_time = 0; s->info.update_time = 0.0; }
    t0 = toc(s);
  }
  s->settings.rho = c_min(c_max(rho_new, OSQP_RHO_MIN), OSQP_RHO_MAX);
  if (s->settings.rho_is_vec) {
    for (i = 0; i < s->m; i++) {
      if (s->constr_type[i] == 0) s->rho_vec[i] = s->settings.rho;
      else if (s->constr_type[i] > 0) s->rho_vec[i] = OSQP_RHO_EQ_OVER_RHO_INEQ * s->settings.rho;
      else s->rho_vec[i] = OSQP_RHO_MIN;
    }
    for (i = 0; i < s->m; i++) s->rho_inv_vec[i] = 1.0 / s->rho_vec[i];
  } else {
    s->rho_inv = 1. / s->settings.rho;
  }
  i = linsys_update_rho(s->linsys, s->settings.rho);
  if (!s->rho_update_from_solve) s->info.update_time += toc(s) - t0;
  return i;
}

/* auxil.c:49-74 */
static int adapt_rho(OracleSolver* s) {
  int flag = 0;
  double rho_new = compute_rho_estimate(s);
  s->info.rho_estimate = rho_new;
  if ((rho_new > s->settings.rho * s->settings.adaptive_rho_tolerance) ||
      (rho_new < s->settings.rho / s->settings.adaptive_rho_tolerance)) {
    flag = oracle_update_rho(s, rho_new);
    s->info.rho_updates += 1;
    s->rho_updated = 1;
  }
  return flag;
}

/* auxil.c:598-674 + scaling.c:194-208 */
static void store_solution(OracleSolver* s, OSQPSolution* sol) {
  int i, n = s->n, m = s->m;
  if (!sol) return;
  if (has_solution(&s->info)) {
    if (s->settings.scaling) {
      for (i = 0; i < n; i++) sol->x[i] = s->x[i] * s->D[i];
      for (i = 0; i < m; i++) { double t = s->y[i] * s->E[i]; sol->y[i] = t * s->cinv; }
    } else {
      memcpy(sol->x, s->x, sizeof(double) * (size_t)n);
      if (m) memcpy(sol->y, s->y, sizeof(double) * (size_t)m);
    }
    for (i = 0; i < m; i++) sol->prim_inf_cert[i] = OSQP_NAN;
    for (i = 0; i < n; i++) sol->dual_inf_cert[i] = OSQP_NAN;
    for (i = 0; i < m; i++) s->delta_y[i] = OSQP_NAN;
    for (i = 0; i < n; i++) s->delta_x[i] = OSQP_NAN;
  } else {
    for (i = 0; i < n; i++) sol->x[i] = OSQP_NAN;
    for (i = 0; i < m; i++) sol->y[i] = OSQP_NAN;
    cold_start(s);
    if (s->info.status_val == OSQP_PRIMAL_INFEASIBLE || s->info.status_val == OSQP_PRIMAL_INFEASIBLE_INACCURATE) {
      double nv = vnorm_inf(s->delta_y, m);
      for (i = 0; i < m; i++) s->delta_y[i] *= 1. / nv;
      memcpy(sol->prim_inf_cert, s->delta_y, sizeof(double) * (size_t)m);
      for (i = 0; i < n; i++) { s->delta_x[i] = OSQP_NAN; sol->dual_inf_cert[i] = OSQP_NAN; }
    }
    if (s->info.status_val == OSQP_DUAL_INFEASIBLE || s->info.status_val == OSQP_DUAL_INFEASIBLE_INACCURATE) {
      double nv = vnorm_inf(s->delta_x, n);
      for (i = 0; i < n; i++) s->delta_x[i] *= 1. / nv;
      memcpy(sol->dual_inf_cert, s->delta_x, sizeof(double) * (size_t)n);
      for (i = 0; i < m; i++) { s->delta_y[i] = OSQP_NAN; sol->prim_inf_cert[i] = OSQP_NAN; }
    }
  }
}

static void swap(double** a, double** b) { double* t = *a; *a = *b; *b = t; }

/* osqp_api.c:636-1038 */
int oracle_solve(OracleSolver* s) {
  int iter, can_check = 0, can_adapt = 0;
  OSQPSettings* st;
  if (!s) return OSQP_WORKSPACE_NOT_INIT_ERROR;
  st = &s->settings;
  if (s->clear_update_time == 1) s->info.update_time = 0.0;
  s->rho_update_from_solve = 1;
  tic(s);
  if (!st->warm_starting) cold_start(s);
  for (iter = 1; iter <= st->max_iter; iter++) {
    double run_time;
    swap(&s->x, &s->x_prev);
    swap(&s->z, &s->z_prev);
    compute_rhs(s);
    linsys_solve(s->linsys, s->xz_tilde, iter);
    update_x(s);
    update_z(s);
    update_y(s);
    run_time = (s->first_run ? s->info.setup_time : s->info.update_time) + toc(s);
    if (st->time_limit && run_time >= st->time_limit) {
      update_status(&s->info, OSQP_TIME_LIMIT_REACHED);
      break;
    }
    can_check = st->check_termination && (iter % st->check_termination == 0);
    switch (st->adaptive_rho) {
      case OSQP_ADAPTIVE_RHO_UPDATE_DISABLED:
        can_adapt = 0;
        break;
      case OSQP_ADAPTIVE_RHO_UPDATE_TIME:
        if (st->adaptive_rho_interval == 0) {
          if (toc(s) > st->adaptive_rho_fraction * s->info.setup_time) {
            int r, base = st->check_termination ? st->check_termination : OSQP_CHECK_TERMINATION;
            r = (int)(floor((double)iter / base + 0.5) * base);
            st->adaptive_rho_interval = c_min(1, c_max(r, st->check_termination));
          } else {
            can_adapt = 0;
            break;
          }
        }
        /* fall through */
      case OSQP_ADAPTIVE_RHO_UPDATE_KKT_ERROR:
      case OSQP_ADAPTIVE_RHO_UPDATE_ITERATIONS:
        can_adapt = (st->adaptive_rho_interval && (iter % st->adaptive_rho_interval == 0)) ? 1 : 0;
        break;
    }
    if (can_check || can_adapt || iter == 1) update_info(s, iter);
    if (can_check && check_termination(s, 0)) break;
    s->rho_updated = 0;
    if (can_adapt && st->adaptive_rho == OSQP_ADAPTIVE_RHO_UPDATE_KKT_ERROR)
      can_adapt = (s->info.rel_kkt_error <= st->adaptive_rho_fraction * s->last_rel_kkt) ? 1 : 0;
    if (can_adapt && adapt_rho(s)) return 1;
    if (s->rho_updated) s->last_rel_kkt = s->info.rel_kkt_error;
  }
  if (!can_check) {
    update_info(s, iter - 1);
    check_termination(s, 0);
  }
  if (has_solution(&s->info))
    compute_obj_val_dual_gap(s, s->x, s->y, &s->info.obj_val, &s->info.dual_obj_val, &s->info.duality_gap);
  if (s->info.status_val == OSQP_UNSOLVED) {
    if (!check_termination(s, 1)) update_status(&s->info, OSQP_MAX_ITER_REACHED);
  }
  if (s->info.status_val == OSQP_TIME_LIMIT_REACHED) {
    if (!check_termination(s, 1)) update_status(&s->info, OSQP_TIME_LIMIT_REACHED);
  }
  s->info.rho_estimate = compute_rho_estimate(s);
  s->info.solve_time = toc(s);
  if (s->first_run) s->info.run_time = s->info.setup_time + s->info.solve_time + s->info.polish_time;
  else s->info.run_time = s->info.update_time + s->info.solve_time + s->info.polish_time;
  if (s->first_run) s->first_run = 0;
  s->clear_update_time = 1;
  s->rho_update_from_solve = 0;
  if (s->has_solution_buf) store_solution(s, &s->solution);
  return 0;
}

/* osqp_api.c:1184-1257 */
int oracle_update_data_vec(OracleSolver* s, const double* q_new, const double* l_new, const double* u_new) {
  int i, n = s->n, m = s->m;
  double t0;
  if (s->clear_update_time == 1) { s->clear_update_time = 0; s->info.update_time = 0.0; }
  t0 = toc(s);
  if (l_new || u_new) {
    double* lt = vnew(m);
    double* ut = vnew(m);
    memcpy(lt, s->l, sizeof(double) * (size_t)m);
    memcpy(ut, s->u, sizeof(double) * (size_t)m);
    if (l_new) for (i = 0; i < m; i++) lt[i] = l_new[i] * (s->settings.scaling ? s->E[i] : 1.0);
    if (u_new) for (i = 0; i < m; i++) ut[i] = u_new[i] * (s->settings.scaling ? s->E[i] : 1.0);
    for (i = 0; i < m; i++)
      if (lt[i] > ut[i]) { free(lt); free(ut); return OSQP_DATA_VALIDATION_ERROR; }
    memcpy(s->l, lt, sizeof(double) * (size_t)m);
    memcpy(s->u, ut, sizeof(double) * (size_t)m);
    free(lt);
    free(ut);
    if (s->settings.rho_is_vec) {
      if (set_rho_vec(s)) linsys_update_rho(s->linsys, s->settings.rho);
    }
  }
  if (q_new) {
    for (i = 0; i < n; i++) s->q[i] = q_new[i];
    if (s->settings.scaling) {
      for (i = 0; i < n; i++) s->q[i] *= s->D[i];
      for (i = 0; i < n; i++) s->q[i] *= s->c;
    }
  }
  reset_info(&s->info);
  s->info.update_time += toc(s) - t0;
  return 0;
}

/* osqp_api.c:1260-1293 */
int oracle_warm_start(OracleSolver* s, const double* x, const double* y) {
  int i;
  if (!s->settings.warm_starting) s->settings.warm_starting = 1;
  if (x) memcpy(s->x, x, sizeof(double) * (size_t)s->n);
  if (y) memcpy(s->y, y, sizeof(double) * (size_t)s->m);
  if (s->settings.scaling) {
    if (x) for (i = 0; i < s->n; i++) s->x[i] *= s->Dinv[i];
    if (y) {
      for (i = 0; i < s->m; i++) s->y[i] *= s->Einv[i];
      for (i = 0; i < s->m; i++) s->y[i] *= s->c;
    }
  }
  if (x) csc_Axpy(s->A, s->x, s->z, 1.0, 0.0);
  linsys_warm_start(s->linsys, s->x);
  return 0;
}

void oracle_cold_start(OracleSolver* s) { cold_start(s); }

/* osqp_api.c:1309-1409 */
int oracle_update_data_mat(OracleSolver* s, const double* Px_new, const int* Px_idx, int P_n,
                           const double* Ax_new, const int* Ax_idx, int A_n) {
  int i, nnzP = ocsc_nnz(s->P), nnzA = ocsc_nnz(s->A), flag;
  double t0;
  if (s->clear_update_time == 1) { s->clear_update_time = 0; s->info.update_time = 0.0; }
  t0 = toc(s);
  if (P_n > nnzP || P_n < 0) return 1;
  if (!Px_idx && P_n != 0 && P_n != nnzP) return 1;
  if (P_n == 0) P_n = nnzP;
  if (A_n > nnzA || A_n < 0) return 2;
  if (!Ax_idx && A_n != 0 && A_n != nnzA) return 2;
  if (A_n == 0) A_n = nnzA;
  if (s->settings.scaling) unscale_data(s);
  if (Px_new) for (i = 0; i < P_n; i++) s->P->x[Px_idx ? Px_idx[i] : i] = Px_new[i];
  if (Ax_new) for (i = 0; i < A_n; i++) s->A->x[Ax_idx ? Ax_idx[i] : i] = Ax_new[i];
  if (s->settings.scaling) scale_data(s);
  flag = linsys_update_matrices(s->linsys);
  reset_info(&s->info);
  s->info.update_time += toc(s) - t0;
  return flag;
}

/* osqp_api.c:1474-1530 */
int oracle_update_settings(OracleSolver* s, const OSQPSettings* ns) {
  OSQPSettings* st = &s->settings;
  if (validate_settings(ns, 0)) return OSQP_SETTINGS_VALIDATION_ERROR;
  st->profiler_level = ns->profiler_level;
  st->verbose = ns->verbose;
  st->warm_starting = ns->warm_starting;
  st->polishing = ns->polishing;
  st->alpha = ns->alpha;
  st->cg_max_iter = ns->cg_max_iter;
  st->cg_tol_reduction = ns->cg_tol_reduction;
  st->cg_tol_fraction = ns->cg_tol_fraction;
  st->cg_precond = ns->cg_precond;
  st->max_iter = ns->max_iter;
  st->eps_abs = ns->eps_abs;
  st->eps_rel = ns->eps_rel;
  st->eps_prim_inf = ns->eps_prim_inf;
  st->eps_dual_inf = ns->eps_dual_inf;
  st->scaled_termination = ns->scaled_termination;
  st->check_termination = ns->check_termination;
  st->check_dualgap = ns->check_dualgap;
  st->time_limit = ns->time_limit;
  st->delta = ns->delta;
  st->polish_refine_iter = ns->polish_refine_iter;
  linsys_update_settings(s->linsys, st);
  return 0;
}

void oracle_cleanup(OracleSolver* s) {
  if (!s) return;
  linsys_free(s->linsys);
  ocsc_free(s->P);
  ocsc_free(s->A);
  free(s->q); free(s->l); free(s->u);
  free(s->rho_vec); free(s->rho_inv_vec); free(s->constr_type);
  free(s->x); free(s->z); free(s->xz_tilde); free(s->x_prev); free(s->z_prev); free(s->y);
  free(s->Ax); free(s->Px); free(s->Aty); free(s->delta_y); free(s->Atdelta_y);
  free(s->delta_x); free(s->Pdelta_x); free(s->Adelta_x);
  free(s->D); free(s->Dinv); free(s->E); free(s->Einv);
  free(s->D_temp); free(s->D_temp_A); free(s->E_temp);
  if (s->has_solution_buf) {
    free(s->solution.x); free(s->solution.y); free(s->solution.prim_inf_cert); free(s->solution.dual_inf_cert);
  }
  free(s);
}

/* accessors for ctypes */
OSQPInfo* oracle_info(OracleSolver* s) { return &s->info; }
OSQPSolution* oracle_solution(OracleSolver* s) { return &s->solution; }
OSQPSettings* oracle_settings(OracleSolver* s) { return &s->settings; }
long oracle_total_pcg_iters(OracleSolver* s) { return s->linsys ? s->linsys->total_pcg_iters : 0; }

/* Scaled problem data and scaling vectors after setup (for bit-wise checks of
 * the device-side Ruiz equilibration). */
void oracle_get_scaling(OracleSolver* s, double* D, double* E, double* c, double* q_scaled,
                        double* l_scaled, double* u_scaled) {
  if (D) memcpy(D, s->D, sizeof(double) * (size_t)s->n);
  if (E && s->m) memcpy(E, s->E, sizeof(double) * (size_t)s->m);
  if (c) *c = s->c;
  if (q_scaled) memcpy(q_scaled, s->q, sizeof(double) * (size_t)s->n);
  if (l_scaled && s->m) memcpy(l_scaled, s->l, sizeof(double) * (size_t)s->m);
  if (u_scaled && s->m) memcpy(u_scaled, s->u, sizeof(double) * (size_t)s->m);
}

void oracle_get_scaled_matrices(OracleSolver* s, double* Px, double* Ax) {
  if (Px) memcpy(Px, s->P->x, sizeof(double) * (size_t)ocsc_nnz(s->P));
  if (Ax) memcpy(Ax, s->A->x, sizeof(double) * (size_t)ocsc_nnz(s->A));
}

/* Direct solve of the reduced KKT operator for linsys-level tests:
 * solves (P + sigma I + A' diag(rho) A) x = b with the same LDL^T machinery
 * (through the full quasi-definite KKT). */
int oracle_kkt_solve(int n, int m, const int* Pp, const int* Pi, const double* Px,
                     const int* Ap, const int* Ai, const double* Ax,
                     double sigma, const double* rho_vec, double* b) {
  ocsc P = ocsc_view(n, n, Pp, Pi, Px), A = ocsc_view(m, n, Ap, Ai, Ax);
  oracle_linsys* s = (oracle_linsys*)calloc(1, sizeof(oracle_linsys));
  int flag;
  s->type = OSQP_DIRECT_SOLVER;
  s->n = n; s->m = m; s->P = &P; s->A = &A; s->rho_vec = rho_vec; s->sigma = sigma; s->rho_sc = 1.0;
  flag = kkt_assemble_factor(s, 1);
  if (!flag) direct_solve(s, b);
  linsys_free(s);
  return flag;
}
