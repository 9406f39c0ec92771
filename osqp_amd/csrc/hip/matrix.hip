// matrix.hip — OSQPMatrix on HBM: CSR of A, CSR of A' (= the input CSC), and
// the full symmetric CSR of P built from its upper triangle.
//
// Reference semantics: algebra/builtin/matrix.c (CSC storage, triu P) and
// algebra/_common/csc_math.c:26-360.  Differences of representation only:
//   * P is stored full (both triangles) so that P x is one gather-SpMV per
//     row.  Scaling keeps the two copies of an off-diagonal entry bit-equal to
//     the triu value the builtin would hold: lmult multiplies entry (r,c) by
//     L[min(r,c)], rmult by R[max(r,c)] — the row/column of the triu entry.
//   * col_norm_inf(P) keeps the builtin meaning (max over the STORED triu
//     column, csc_math.c:292-308 via matrix.c OSQPMatrix_col_norm_inf), i.e.
//     row r of the full CSR restricted to columns <= r.
// Structure work (CSC -> CSR, triu -> full, slot maps) is integer-exact: a
// stable counting sort on the host, the same ordering the oracle's
// oracle_csc_to_csr / oracle_triu_to_full_csr produce.
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "hip_internal.h"
#include "panel.h"
#include "spmv_rb.h"

using namespace osqp_hip;

namespace {

inline hipStream_t S() { return rt().stream; }

#define GRID_STRIDE(i, n) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

inline int egrid(long long n) { return grid_for(n, kBlock, 16384); }

__global__ void k_scale(double* v, double sc, long long n) { GRID_STRIDE(i, n) v[i] *= sc; }

// per-entry row index of a CSR (expanded once per call; rows are short)
// val[k] *= L[r] over the entries of row r; one wavefront per row, so the
// row's entries are read and written in coalesced strides
__global__ void k_lmult_csr(double* val, const int* rp, const double* L, int rows) {
  const int lane = threadIdx.x & (kWave - 1);
  const long long nw = (long long)gridDim.x * blockDim.x / kWave;
  for (long long r = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave; r < rows; r += nw) {
    const double d = L[r];
    const int e = rp[r + 1];
    for (int k = rp[r] + lane; k < e; k += kWave) val[k] *= d;
  }
}

__global__ void k_rmult_csr(double* val, const int* ci, const double* R, long long nnz) {
  GRID_STRIDE(k, nnz) val[k] *= R[ci[k]];
}

// P full: entry (r,c) scaled by the factor of the triu entry it copies
__global__ void k_lmult_sym(double* val, const int* rp, const int* ci, const double* L, int rows) {
  GRID_STRIDE(r, rows) {
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      int c = ci[k];
      val[k] *= L[r < c ? r : c];
    }
  }
}

__global__ void k_rmult_sym(double* val, const int* rp, const int* ci, const double* R, int rows) {
  GRID_STRIDE(r, rows) {
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      int c = ci[k];
      val[k] *= R[r > c ? r : c];
    }
  }
}

// y = alpha * M x + beta * y over nonzero-balanced row blocks
__global__ __launch_bounds__(kBlock) void k_rb_spmv(const int* __restrict__ rows, int nb, const int* __restrict__ rp,
                                                   const int* __restrict__ ci, const double* __restrict__ val,
                                                   const double* __restrict__ x, double* __restrict__ y, double alpha,
                                                   double beta) {
  __shared__ __attribute__((aligned(16))) double prod[kChunk];
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    rb_spmv_block<double, double>(rows[b], rows[b + 1], rp, ci, val, prod, [&](int c) { return x[c]; },
                                  [&](int row, double sum) {
                                    double v = alpha * sum;
                                    if (beta != 0.0) v += beta * y[row];
                                    y[row] = v;
                                  });
  }
}

__global__ void k_scale_y(double* y, double beta, int n) {
  GRID_STRIDE(i, n) y[i] = (beta == 0.0) ? 0.0 : beta * y[i];
}

// max |M_rc| over row r, optionally restricted to columns c <= r
template <int G>
__global__ __launch_bounds__(kBlock) void k_row_max(int rows, const int* rp, const int* ci, const double* val,
                                                   double* out, int lower_only) {
  const int lane = threadIdx.x & (G - 1);
  const int rows_per_block = kBlock / G;
  for (int row = blockIdx.x * rows_per_block + threadIdx.x / G; row < rows; row += gridDim.x * rows_per_block) {
    double acc = 0.0;
    const int e = rp[row + 1];
    for (int k = rp[row] + lane; k < e; k += G) {
      if (lower_only && ci[k] > row) continue;
      double a = fabs(val[k]);
      acc = a > acc ? a : acc;
    }
    acc = group_max<G>(acc);
    if (lane == 0) out[row] = acc;
  }
}

// d_r = sum_k val_k^2 * D[ci_k] over row r (diag of A' D A from the CSR of A')
template <int G>
__global__ __launch_bounds__(kBlock) void k_row_sq_weighted(int rows, const int* rp, const int* ci, const double* val,
                                                           const double* D, double* out) {
  const int lane = threadIdx.x & (G - 1);
  const int rows_per_block = kBlock / G;
  for (int row = blockIdx.x * rows_per_block + threadIdx.x / G; row < rows; row += gridDim.x * rows_per_block) {
    double acc = 0.0;
    const int e = rp[row + 1];
    for (int k = rp[row] + lane; k < e; k += G) acc += val[k] * val[k] * D[ci[k]];
    acc = group_sum<G>(acc);
    if (lane == 0) out[row] = acc;
  }
}

__global__ void k_extract_diag(const double* val, const int* diag_pos, double* d, int n) {
  GRID_STRIDE(i, n) d[i] = diag_pos[i] >= 0 ? val[diag_pos[i]] : 0.0;
}

// scatter new CSC values into the CSR copy (+ mirror) and the transpose copy
__global__ void k_update_vals(const double* xnew, const int* idx, long long cnt, const int* csc2csr,
                              const int* csc2mirror, double* val, double* tval) {
  GRID_STRIDE(t, cnt) {
    int k = idx ? idx[t] : (int)t;
    double v = xnew[t];
    val[csc2csr[k]] = v;
    if (csc2mirror && csc2mirror[k] >= 0) val[csc2mirror[k]] = v;
    if (tval) tval[k] = v;
  }
}

template <template <int> class K, typename... Args>
void launch_rows(int G, int rows, Args... args) {
  if (rows <= 0) return;
  const int rpb = kBlock / G;
  int grid = (int)((rows + rpb - 1) / rpb);
  if (grid > kRowGridCap) grid = kRowGridCap;
  switch (G) {
    case 4: hipLaunchKernelGGL(K<4>::fn, dim3(grid), dim3(kBlock), 0, S(), args...); break;
    case 8: hipLaunchKernelGGL(K<8>::fn, dim3(grid), dim3(kBlock), 0, S(), args...); break;
    case 16: hipLaunchKernelGGL(K<16>::fn, dim3(grid), dim3(kBlock), 0, S(), args...); break;
    case 32: hipLaunchKernelGGL(K<32>::fn, dim3(grid), dim3(kBlock), 0, S(), args...); break;
    default: hipLaunchKernelGGL(K<64>::fn, dim3(grid), dim3(kBlock), 0, S(), args...); break;
  }
  LAUNCHCHK();
}

template <int G> struct RowMaxK { static constexpr auto fn = k_row_max<G>; };
template <int G> struct RowSqK { static constexpr auto fn = k_row_sq_weighted<G>; };

template <typename T>
T* dev_copy(const T* h, size_t n) {
  T* d = (T*)dev_alloc(sizeof(T) * (n > 0 ? n : 1));
  if (d && n) upload(d, h, sizeof(T) * n);
  return d;
}

void spmv(const int* rp, const int* ci, const double* val, int rows, long long nnz, const int* rb_rows, int rb_n,
          const double* x, double* y, double alpha, double beta) {
  if (rows <= 0) return;
  if (nnz == 0 || alpha == 0.0) {
    hipLaunchKernelGGL(k_scale_y, dim3(egrid(rows)), dim3(kBlock), 0, S(), y, beta, rows);
    return;
  }
  int grid = rb_n < kRowGridCap ? rb_n : kRowGridCap;
  hipLaunchKernelGGL(k_rb_spmv, dim3(grid), dim3(kBlock), 0, S(), rb_rows, rb_n, rp, ci, val, x, y, alpha, beta);
  LAUNCHCHK();
}

// y = alpha * (sum of a row's panel partials) + beta * y (the row-block
// kernel's epilogue)
template <int G>
__global__ void k_pan_axpy(int R, int np, const double* partial, double alpha, double beta, double* y) {
  COMB_ROWS(i, j, R, G) {
    const double s = panel_sum<double, G>(partial, np, R, i, j);
    if (j == 0) {
      double v = alpha * s;
      if (beta != 0.0) v += beta * y[i];
      y[i] = v;
    }
  }
}

// Products of matrices with at least this many nonzeros go through the
// matrix's FP64 panels (built with the matrix); smaller ones stay on the
// one-launch row-block kernel, where launch latency, not traffic, is the cost.
constexpr long long kPanelAxpyMinNnz = 1LL << 20;

// FP64 panels of M (t = 0) or M' (t = 1) from the host CSR structure of that
// operator; marks the slot tried whatever the outcome.
void build_matrix_panels(OSQPMatrix* M, int t, const int* rp, const int* ci) {
  const int R = t ? M->n : M->m, C = t ? M->m : M->n;
  M->pan_state[t] = -1;
  if (M->is_triu || R <= 0 || C <= 0 || M->nnz <= 0 || !panel_worth(M->nnz, R, C, 64)) return;
  PanelMatrix* P = (PanelMatrix*)calloc(1, sizeof(PanelMatrix));
  if (!P) return;
  if (panel_build(P, R, C, 64, rp, ci)) {
    panel_free(P);
    free(P);
    return;
  }
  const size_t need = (size_t)P->np * (size_t)R;
  if (need > M->pan_part_len) {
    double* b = (double*)dev_alloc(sizeof(double) * need);
    if (!b) {
      panel_free(P);
      free(P);
      return;
    }
    dev_free(M->pan_part);
    M->pan_part = b;
    M->pan_part_len = need;
  }
  M->pan64[t] = P;
  M->pan_state[t] = 1;
}

// y = alpha M x + beta y (t = 0) or alpha M' x + beta y (t = 1) over the
// panels; false when the product stays on the row-block kernel
bool panel_axpy(const OSQPMatrix* M, int t, const double* x, double* y, double alpha, double beta) {
  if (M->is_triu || !x || alpha == 0.0 || M->nnz < kPanelAxpyMinNnz) return false;
  const PanelMatrix* P = matrix_panels64(M, t);
  if (!P || P->ntask <= 0) return false;
  panel_spmv(*P, nullptr, x, P->C, (const double*)nullptr, M->pan_part);
  const int R = P->R, G = comb_group(P->np, R);
  const dim3 g(grid_for((long long)R * G, kBlock, 8192));
  switch (G) {
    case 8: hipLaunchKernelGGL(k_pan_axpy<8>, g, dim3(kBlock), 0, S(), R, P->np, (const double*)M->pan_part, alpha, beta, y); break;
    case 4: hipLaunchKernelGGL(k_pan_axpy<4>, g, dim3(kBlock), 0, S(), R, P->np, (const double*)M->pan_part, alpha, beta, y); break;
    case 2: hipLaunchKernelGGL(k_pan_axpy<2>, g, dim3(kBlock), 0, S(), R, P->np, (const double*)M->pan_part, alpha, beta, y); break;
    default: hipLaunchKernelGGL(k_pan_axpy<1>, g, dim3(kBlock), 0, S(), R, P->np, (const double*)M->pan_part, alpha, beta, y); break;
  }
  LAUNCHCHK();
  return true;
}

}  // namespace

namespace osqp_hip {

const PanelMatrix* matrix_panels64(const OSQPMatrix* Mc, int t) {
  OSQPMatrix* M = (OSQPMatrix*)Mc;
  if (M->is_triu) return nullptr;
  if (M->pan_state[t] == 0) {
    const int R = t ? M->n : M->m;
    std::vector<int> hrp((size_t)R + 1), hci((size_t)(M->nnz > 0 ? M->nnz : 1));
    download(hrp.data(), t ? M->trp : M->rp, sizeof(int) * hrp.size());
    if (M->nnz) download(hci.data(), t ? M->tci : M->ci, sizeof(int) * (size_t)M->nnz);
    build_matrix_panels(M, t, hrp.data(), hci.data());
  }
  PanelMatrix* P = M->pan64[t];
  if (P && P->ver != M->version) {
    panel_refresh_values(P, t ? M->tval : M->val);
    P->ver = M->version;
  }
  return P;
}

}  // namespace osqp_hip

// ---- host-side structure builders (also used by the batched engine) ----------
namespace osqp_hip {

int build_row_blocks(const int* rp, int rows, int* out) {
  int nb = 0, start = 0;
  out[0] = 0;
  for (int r = 0; r < rows; ++r) {
    const int len = rp[r + 1] - rp[r];
    if (len > kCap) {
      if (start < r) out[++nb] = r;      // close the open block
      out[++nb] = r + 1;                 // the long row alone
      start = r + 1;
    } else if (rp[r + 1] - rp[start] > kCap || r - start >= kChunk) {
      out[++nb] = r;                     // block [start, r)
      start = r;
    }
  }
  if (start < rows) out[++nb] = rows;
  return nb;
}

static int* upload_row_blocks(const int* rp_host, int rows, int* nb_out) {
  std::vector<int> b((size_t)rows + 2);
  int nb = build_row_blocks(rp_host, rows, b.data());
  *nb_out = nb;
  int* d = (int*)dev_alloc(sizeof(int) * (size_t)(nb + 1));
  upload(d, b.data(), sizeof(int) * (size_t)(nb + 1));
  return d;
}

// CSR of an m x n CSC matrix by a stable counting sort; csc2csr[k] = CSR slot
void build_csr_from_csc(int m, int n, const int* p, const int* i, std::vector<int>& rp, std::vector<int>& ci,
                        std::vector<int>& csc2csr) {
  const int nnz = p[n];
  rp.assign((size_t)m + 1, 0);
  ci.resize((size_t)nnz);
  csc2csr.resize((size_t)nnz);
  for (int k = 0; k < nnz; ++k) rp[(size_t)i[k] + 1]++;
  for (int r = 0; r < m; ++r) rp[(size_t)r + 1] += rp[(size_t)r];
  std::vector<int> next(rp.begin(), rp.end() - 1);
  for (int j = 0; j < n; ++j)
    for (int k = p[j]; k < p[j + 1]; ++k) {
      int pos = next[(size_t)i[k]]++;
      ci[(size_t)pos] = j;
      csc2csr[(size_t)k] = pos;
    }
}

// Full symmetric CSR of an n x n matrix from its upper-triangular CSC.
// Row r lists (r,c) for c<r (mirrors of triu column r) then c>=r, sorted.
void build_full_from_triu(int n, const int* p, const int* i, std::vector<int>& rp, std::vector<int>& ci,
                          std::vector<int>& pos, std::vector<int>& mirror, std::vector<int>& diag) {
  const int nnz = p[n];
  rp.assign((size_t)n + 1, 0);
  for (int j = 0; j < n; ++j)
    for (int k = p[j]; k < p[j + 1]; ++k) {
      rp[(size_t)i[k] + 1]++;
      if (i[k] != j) rp[(size_t)j + 1]++;
    }
  for (int r = 0; r < n; ++r) rp[(size_t)r + 1] += rp[(size_t)r];
  const int full = rp[(size_t)n];
  ci.resize((size_t)full);
  pos.resize((size_t)nnz);
  mirror.resize((size_t)nnz);
  diag.assign((size_t)n, -1);
  std::vector<int> next(rp.begin(), rp.end() - 1);
  for (int j = 0; j < n; ++j) {
    for (int k = p[j]; k < p[j + 1]; ++k) {
      int row = i[k];
      if (row == j) continue;
      int q = next[(size_t)j]++;
      ci[(size_t)q] = row;
      mirror[(size_t)k] = q;
    }
    for (int k = p[j]; k < p[j + 1]; ++k) {
      int row = i[k];
      int q = next[(size_t)row]++;
      ci[(size_t)q] = j;
      pos[(size_t)k] = q;
      if (row == j) {
        mirror[(size_t)k] = -1;
        diag[(size_t)j] = q;
      }
    }
  }
}

}  // namespace osqp_hip

extern "C" {

OSQPInt osqp_amd_matrix_export(const OSQPMatrix* M, OSQPInt* nnz_csr, OSQPInt* nnz_in, OSQPInt* rp, OSQPInt* ci,
                               OSQPFloat* val, OSQPInt* trp, OSQPInt* tci, OSQPFloat* tval, OSQPInt* csc2csr,
                               OSQPInt* csc2mirror, OSQPInt* diag) {
  if (!M) return 1;
  const int rows = M->is_triu ? M->n : M->m;
  if (nnz_csr) *nnz_csr = M->nnz;
  if (nnz_in) *nnz_in = M->nnz_in;
  if (rp) download(rp, M->rp, sizeof(int) * ((size_t)rows + 1));
  if (ci && M->nnz) download(ci, M->ci, sizeof(int) * (size_t)M->nnz);
  if (val && M->nnz) download(val, M->val, sizeof(double) * (size_t)M->nnz);
  if (!M->is_triu) {
    if (trp) download(trp, M->trp, sizeof(int) * ((size_t)M->n + 1));
    if (tci && M->nnz_in) download(tci, M->tci, sizeof(int) * (size_t)M->nnz_in);
    if (tval && M->nnz_in) download(tval, M->tval, sizeof(double) * (size_t)M->nnz_in);
  }
  if (csc2csr && M->nnz_in) download(csc2csr, M->csc2csr, sizeof(int) * (size_t)M->nnz_in);
  if (M->is_triu) {
    if (csc2mirror && M->nnz_in) download(csc2mirror, M->csc2mirror, sizeof(int) * (size_t)M->nnz_in);
    if (diag && M->n) download(diag, M->diag_pos, sizeof(int) * (size_t)M->n);
  }
  return 0;
}

OSQPMatrix* OSQPMatrix_new_from_csc(const OSQPCscMatrix* A, OSQPInt is_triu) {
  OSQPMatrix* M = (OSQPMatrix*)calloc(1, sizeof(OSQPMatrix));
  if (!M) return nullptr;
  const int m = A->m, n = A->n, nnz = A->p[n];
  M->m = m;
  M->n = n;
  M->is_triu = is_triu ? 1 : 0;
  M->nnz_in = nnz;
  std::vector<int> rp, ci, pos, mirror, diag;
  if (!is_triu) {
    build_csr_from_csc(m, n, A->p, A->i, rp, ci, pos);
    std::vector<double> val((size_t)(nnz > 0 ? nnz : 1));
    for (int k = 0; k < nnz; ++k) val[(size_t)pos[(size_t)k]] = A->x[k];
    M->nnz = nnz;
    M->rp = dev_copy(rp.data(), rp.size());
    M->ci = dev_copy(ci.data(), ci.size());
    M->val = dev_copy(val.data(), (size_t)nnz);
    M->trp = dev_copy(A->p, (size_t)n + 1);
    M->tci = dev_copy(A->i, (size_t)nnz);
    M->tval = dev_copy(A->x, (size_t)nnz);
    M->csc2csr = dev_copy(pos.data(), pos.size());
    M->csc2mirror = nullptr;
    M->diag_pos = nullptr;
    M->lanes = lanes_for_rows(nnz, m);
    M->tlanes = lanes_for_rows(nnz, n);
    M->rb_rows = upload_row_blocks(rp.data(), m, &M->rb_n);
    M->trb_rows = upload_row_blocks(A->p, n, &M->trb_n);
    // large operators: FP64 panels of A and A' now, from the host structure
    // (values follow on first use)
    if (nnz >= kPanelAxpyMinNnz) {
      build_matrix_panels(M, 0, rp.data(), ci.data());
      build_matrix_panels(M, 1, A->p, A->i);
    }
  } else {
    build_full_from_triu(n, A->p, A->i, rp, ci, pos, mirror, diag);
    const int full = rp[(size_t)n];
    std::vector<double> val((size_t)(full > 0 ? full : 1));
    for (int k = 0; k < nnz; ++k) {
      val[(size_t)pos[(size_t)k]] = A->x[k];
      if (mirror[(size_t)k] >= 0) val[(size_t)mirror[(size_t)k]] = A->x[k];
    }
    M->nnz = full;
    M->rp = dev_copy(rp.data(), rp.size());
    M->ci = dev_copy(ci.data(), ci.size());
    M->val = dev_copy(val.data(), (size_t)full);
    M->csc2csr = dev_copy(pos.data(), pos.size());
    M->csc2mirror = dev_copy(mirror.data(), mirror.size());
    M->diag_pos = dev_copy(diag.data(), diag.size());
    M->lanes = M->tlanes = lanes_for_rows(full, n);
    M->rb_rows = upload_row_blocks(rp.data(), n, &M->rb_n);
    M->trb_rows = nullptr;
    M->trb_n = 0;
  }
  M->fp32_dirty = 1;
  M->version = 1;
  return M;
}

void OSQPMatrix_free(OSQPMatrix* M) {
  if (!M) return;
  dev_free(M->rp); dev_free(M->ci); dev_free(M->val);
  dev_free(M->trp); dev_free(M->tci); dev_free(M->tval);
  dev_free(M->csc2csr); dev_free(M->csc2mirror); dev_free(M->diag_pos);
  dev_free(M->val32); dev_free(M->tval32);
  dev_free(M->rb_rows); dev_free(M->trb_rows);
  for (int t = 0; t < 2; ++t)
    if (M->pan64[t]) {
      panel_free(M->pan64[t]);
      free(M->pan64[t]);
    }
  dev_free(M->pan_part);
  free(M);
}

OSQPInt OSQPMatrix_get_m(const OSQPMatrix* M) { return M->m; }
OSQPInt OSQPMatrix_get_n(const OSQPMatrix* M) { return M->n; }
OSQPInt OSQPMatrix_get_nz(const OSQPMatrix* M) { return M->nnz_in; }

static void touched(OSQPMatrix* M) {
  M->fp32_dirty = 1;
  M->version++;
}

void OSQPMatrix_update_values(OSQPMatrix* M, const OSQPFloat* Mx_new, const OSQPInt* Mx_new_idx, OSQPInt M_new_n) {
  if (M_new_n <= 0 || !Mx_new) return;
  if (M_new_n > M->nnz_in) return;
  // indices index the input CSC; an out-of-range one would be a device write
  // out of bounds, so they are checked here on the host
  if (Mx_new_idx)
    for (OSQPInt t = 0; t < M_new_n; ++t)
      if (Mx_new_idx[t] < 0 || Mx_new_idx[t] >= M->nnz_in) {
        fprintf(stderr, "OSQPMatrix_update_values: index %d out of range [0, %d)\n", (int)Mx_new_idx[t],
                (int)M->nnz_in);
        return;
      }
  double* dx = dev_copy(Mx_new, (size_t)M_new_n);
  int* di = Mx_new_idx ? dev_copy(Mx_new_idx, (size_t)M_new_n) : nullptr;
  hipLaunchKernelGGL(k_update_vals, dim3(egrid(M_new_n)), dim3(kBlock), 0, S(), dx, di, (long long)M_new_n,
                     M->csc2csr, M->csc2mirror, M->val, M->is_triu ? nullptr : M->tval);
  LAUNCHCHK();
  dev_free(dx);
  dev_free(di);
  touched(M);
}

void OSQPMatrix_mult_scalar(OSQPMatrix* A, OSQPFloat sc) {
  if (A->nnz > 0) hipLaunchKernelGGL(k_scale, dim3(egrid(A->nnz)), dim3(kBlock), 0, S(), A->val, sc, (long long)A->nnz);
  if (!A->is_triu && A->nnz_in > 0)
    hipLaunchKernelGGL(k_scale, dim3(egrid(A->nnz_in)), dim3(kBlock), 0, S(), A->tval, sc, (long long)A->nnz_in);
  touched(A);
}

void OSQPMatrix_lmult_diag(OSQPMatrix* A, const OSQPVectorf* L) {
  if (A->is_triu) {
    hipLaunchKernelGGL(k_lmult_sym, dim3(egrid(A->n)), dim3(kBlock), 0, S(), A->val, A->rp, A->ci, L->d, A->n);
  } else {
    if (A->m > 0)
      hipLaunchKernelGGL(k_lmult_csr, dim3(grid_for((long long)A->m * kWave, kBlock, 16384)), dim3(kBlock), 0, S(),
                         A->val, A->rp, L->d, A->m);
    if (A->nnz_in > 0)
      hipLaunchKernelGGL(k_rmult_csr, dim3(egrid(A->nnz_in)), dim3(kBlock), 0, S(), A->tval, A->tci, L->d,
                         (long long)A->nnz_in);
  }
  touched(A);
}

void OSQPMatrix_rmult_diag(OSQPMatrix* A, const OSQPVectorf* R) {
  if (A->is_triu) {
    hipLaunchKernelGGL(k_rmult_sym, dim3(egrid(A->n)), dim3(kBlock), 0, S(), A->val, A->rp, A->ci, R->d, A->n);
  } else {
    if (A->nnz > 0)
      hipLaunchKernelGGL(k_rmult_csr, dim3(egrid(A->nnz)), dim3(kBlock), 0, S(), A->val, A->ci, R->d, (long long)A->nnz);
    if (A->n > 0)
      hipLaunchKernelGGL(k_lmult_csr, dim3(grid_for((long long)A->n * kWave, kBlock, 16384)), dim3(kBlock), 0, S(),
                         A->tval, A->trp, R->d, A->n);
  }
  touched(A);
}

void OSQPMatrix_Axpy(const OSQPMatrix* A, const OSQPVectorf* x, OSQPVectorf* y, OSQPFloat alpha, OSQPFloat beta) {
  int rows = A->is_triu ? A->n : A->m;
  if (panel_axpy(A, 0, x ? x->d : nullptr, y->d, alpha, beta)) return;
  spmv(A->rp, A->ci, A->val, rows, A->nnz, A->rb_rows, A->rb_n, x ? x->d : nullptr, y->d, alpha, beta);
}

void OSQPMatrix_Atxpy(const OSQPMatrix* A, const OSQPVectorf* x, OSQPVectorf* y, OSQPFloat alpha, OSQPFloat beta) {
  if (A->is_triu) {
    spmv(A->rp, A->ci, A->val, A->n, A->nnz, A->rb_rows, A->rb_n, x->d, y->d, alpha, beta);
  } else {
    if (panel_axpy(A, 1, x ? x->d : nullptr, y->d, alpha, beta)) return;
    spmv(A->trp, A->tci, A->tval, A->n, A->nnz_in, A->trb_rows, A->trb_n, x ? x->d : nullptr, y->d, alpha, beta);
  }
}

double osqp_amd_matrix_product_bytes(const OSQPMatrix* M, OSQPInt transpose) {
  if (!M) return 0.0;
  const int t = (transpose && !M->is_triu) ? 1 : 0;
  const double rows = M->is_triu ? M->n : (t ? M->n : M->m), cols = M->is_triu ? M->n : (t ? M->m : M->n);
  // the panel path of panel_axpy: the panel SpMV, its partials read once by
  // the combine, y written
  if (!M->is_triu && M->nnz >= kPanelAxpyMinNnz && M->pan_state[t] == 1 && M->pan64[t] &&
      M->pan64[t]->ntask > 0) {
    const PanelMatrix& P = *M->pan64[t];
    return panel_spmv_bytes(P, 8) + (double)P.np * P.R * 8.0 + rows * 8.0;
  }
  // row-block CSR: values, 32-bit indices, row pointers, x once, y written
  const double nnz = t ? (double)M->nnz_in : (double)M->nnz;
  return nnz * 12.0 + (rows + 1.0) * 4.0 + cols * 8.0 + rows * 8.0;
}

void OSQPMatrix_col_norm_inf(const OSQPMatrix* M, OSQPVectorf* E) {
  if (M->is_triu) launch_rows<RowMaxK>(M->lanes, M->n, M->n, (const int*)M->rp, (const int*)M->ci, (const double*)M->val, E->d, 1);
  else launch_rows<RowMaxK>(M->tlanes, M->n, M->n, (const int*)M->trp, (const int*)M->tci, (const double*)M->tval, E->d, 0);
}

void OSQPMatrix_row_norm_inf(const OSQPMatrix* M, OSQPVectorf* E) {
  int rows = M->is_triu ? M->n : M->m;
  launch_rows<RowMaxK>(M->lanes, rows, rows, (const int*)M->rp, (const int*)M->ci, (const double*)M->val, E->d, 0);
}

void OSQPMatrix_AtDA_extract_diag(const OSQPMatrix* A, const OSQPVectorf* D, OSQPVectorf* d) {
  if (A->is_triu) return;
  launch_rows<RowSqK>(A->tlanes, A->n, A->n, (const int*)A->trp, (const int*)A->tci, (const double*)A->tval,
                      (const double*)D->d, d->d);
}

void OSQPMatrix_extract_diag(const OSQPMatrix* A, OSQPVectorf* d) {
  if (!A->is_triu || A->n <= 0) return;
  hipLaunchKernelGGL(k_extract_diag, dim3(egrid(A->n)), dim3(kBlock), 0, S(), A->val, A->diag_pos, d->d, A->n);
}

OSQPInt OSQPMatrix_is_eq(const OSQPMatrix* A, const OSQPMatrix* B, OSQPFloat tol) {
  if (A->m != B->m || A->n != B->n || A->is_triu != B->is_triu || A->nnz != B->nnz) return 0;
  int rows = A->is_triu ? A->n : A->m;
  std::vector<int> ra((size_t)rows + 1), rb((size_t)rows + 1), ca((size_t)A->nnz), cb((size_t)A->nnz);
  std::vector<double> va((size_t)A->nnz), vb((size_t)A->nnz);
  download(ra.data(), A->rp, sizeof(int) * ra.size());
  download(rb.data(), B->rp, sizeof(int) * rb.size());
  if (A->nnz) {
    download(ca.data(), A->ci, sizeof(int) * ca.size());
    download(cb.data(), B->ci, sizeof(int) * cb.size());
    download(va.data(), A->val, sizeof(double) * va.size());
    download(vb.data(), B->val, sizeof(double) * vb.size());
  }
  if (ra != rb || ca != cb) return 0;
  for (size_t k = 0; k < va.size(); ++k)
    if (fabs(va[k] - vb[k]) > tol) return 0;
  return 1;
}

// Rows of A selected by rows[i] != 0, in order (polish's Ared).  Built on the
// host from the device CSR; integer-exact row compaction.
OSQPMatrix* OSQPMatrix_submatrix_byrows(const OSQPMatrix* A, const OSQPVectori* rows) {
  if (A->is_triu) return nullptr;
  const int m = A->m, n = A->n;
  std::vector<int> sel((size_t)(m > 0 ? m : 1)), rp((size_t)m + 1), ci((size_t)(A->nnz > 0 ? A->nnz : 1));
  std::vector<double> val((size_t)(A->nnz > 0 ? A->nnz : 1));
  if (m) download(sel.data(), rows->d, sizeof(int) * (size_t)m);
  download(rp.data(), A->rp, sizeof(int) * rp.size());
  if (A->nnz) {
    download(ci.data(), A->ci, sizeof(int) * (size_t)A->nnz);
    download(val.data(), A->val, sizeof(double) * (size_t)A->nnz);
  }
  // CSC of the selected rows (column-major, rows renumbered)
  std::vector<int> newrow((size_t)(m > 0 ? m : 1), -1);
  int mr = 0;
  for (int r = 0; r < m; ++r)
    if (sel[(size_t)r]) newrow[(size_t)r] = mr++;
  std::vector<int> cp((size_t)n + 1, 0);
  for (int r = 0; r < m; ++r)
    if (newrow[(size_t)r] >= 0)
      for (int k = rp[(size_t)r]; k < rp[(size_t)r + 1]; ++k) cp[(size_t)ci[(size_t)k] + 1]++;
  for (int j = 0; j < n; ++j) cp[(size_t)j + 1] += cp[(size_t)j];
  const int nz = cp[(size_t)n];
  std::vector<int> ri((size_t)(nz > 0 ? nz : 1)), nxt(cp.begin(), cp.end() - 1);
  std::vector<double> rx((size_t)(nz > 0 ? nz : 1));
  for (int r = 0; r < m; ++r)
    if (newrow[(size_t)r] >= 0)
      for (int k = rp[(size_t)r]; k < rp[(size_t)r + 1]; ++k) {
        int q = nxt[(size_t)ci[(size_t)k]]++;
        ri[(size_t)q] = newrow[(size_t)r];
        rx[(size_t)q] = val[(size_t)k];
      }
  OSQPCscMatrix C;
  C.m = mr; C.n = n; C.p = cp.data(); C.i = ri.data(); C.x = rx.data(); C.nzmax = nz; C.nz = -1; C.owned = 0;
  return OSQPMatrix_new_from_csc(&C, 0);
}

}  // extern "C"
