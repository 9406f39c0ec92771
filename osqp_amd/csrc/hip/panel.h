// panel.h — LDS-staged, column-panelled SpMV for large sparse operators.
//
// Random gathers of a multiplied vector larger than a CU's L1 go to L2 one
// request per lane, and that request rate — not HBM bandwidth — bounds a CSR
// SpMV on gfx950 (a 500000 x 50000 matrix with 100 nonzeros per row streams
// at 5.4 TB/s without the gather and at 1.4 TB/s with it).  So a CSR matrix
// (R rows, C columns) is re-stored as np = ceil(C / W) column panels, each a
// CSR over all R rows with 16-bit panel-local column indices, and the SpMV
// gathers from a copy of the panel's slice of x in LDS:
//
//   * work items ("tasks") are row ranges of one panel holding at most
//     64 * PER - (V - 1) nonzeros (V = 4 for FP32, 2 for FP64) and at most 63
//     rows, listed panel by panel (a panel row longer than that is a task of
//     its own, reduced in several passes);
//     the task descriptor carries the nonzero range so a wavefront issues its
//     stream loads without waiting for a row pointer;
//   * a workgroup walks a contiguous range of tasks; for every panel in the
//     range it loads the x slice into LDS once (all loads in flight, then the
//     LDS stores) and its wavefronts take the panel's tasks independently —
//     no workgroup barrier inside a panel;
//   * a wavefront streams the window of 64 * PER (index, value) pairs from
//     the task's first nonzero rounded down to V, each lane V consecutive
//     pairs per load (a 16-byte value load: the address unit's cost is per
//     instruction), with branch-free loads (arrays are padded by 64 * PER
//     entries); it multiplies with the LDS slice into a per-wavefront LDS
//     product buffer (zero outside the task), then reduces each row in
//     column order (groups of L lanes per row, L a power of two chosen from the
//     task's row count; in-order sums within a lane, then a fixed butterfly);
//   * row totals go to partial[p * R + row]; combine kernels add the np
//     partials of a row in panel order.  Results are deterministic and do not
//     depend on the grid size.
#pragma once

#include <stdint.h>

#include "hip_internal.h"

namespace osqp_hip {

constexpr int kPanelRowCap = 63;  // rows per task: one row pointer per lane

struct PanelMatrix {
  int R, C;
  int W;    // columns per panel
  int per;  // nonzeros per lane per task window (task holds <= 64 * per - (V - 1))
  int np;
  long long nnz;
  int* rp;         // np * (R + 1) row pointers, absolute into ci/val
  uint16_t* ci;    // nnz + 64 * per entries: column index within the panel
  int* src;        // nnz: CSR slot each panel entry copies its value from
  int ntask;
  int4* task;      // ntask: first row, end row, first nonzero, end nonzero
  int2* tpanel;    // ntask: panel of each task, end of that panel's tasks
  float* val32;    // nnz + pad, optional
  double* val64;   // nnz + pad, optional
  long long ver;   // source-value version the copies were made from
  int grid;        // workgroups per launch
  int* pcol;       // np: column block (of W columns) of each panel; null when every block is a panel
};

// Host builder from a host CSR structure (R x C): allocates and uploads the
// structure for the FP32 (bits 32) or FP64 (bits 64) kernel; values are
// filled by panel_refresh_values.  With `sel` (one flag per block of W
// columns), only the selected blocks become panels and the entries of the
// others are left out (the caller handles them).  Returns 0 on success.
int panel_build(PanelMatrix* P, int R, int C, int bits, const int* rp, const int* ci,
                const unsigned char* sel = nullptr);
// columns per panel for the FP32 (bits 32) and FP64 (bits 64) kernels
int panel_width(int bits);
void panel_free(PanelMatrix* P);
// Whether panels pay for an R x C operator with nnz nonzeros: one panel, or
// at least 4 nonzeros per (panel, row) on average (the row pointers and
// partial sums cost about as much as 2 nonzeros per (panel, row)).
int panel_worth(long long nnz, int R, int C, int bits);
// val64 or val32 <- csr_val[src] (device array, FP64), on the backend stream.
void panel_refresh_values(PanelMatrix* P, const double* csr_val);
// partial[p * R + row] = sum over panel p of row `row`, for x = [x0 (n0); x1].
// When `active` is non-null the launch does nothing while *active == 0.
void panel_spmv(const PanelMatrix& P, const int* active, const float* x0, int n0, const float* x1, float* partial);
void panel_spmv(const PanelMatrix& P, const int* active, const double* x0, int n0, const double* x1,
                double* partial);
// The same product over FP64-built panels (4096 columns, 6 per lane) with
// FP32 copies of their values: x, the products and the sums in FP64.  val32
// comes from panel_alloc_f32vals (nnz plus the zeroed padding) and
// panel_fill_f32vals (val32[k] = (float)csr_val[src[k]]).
void panel_spmv_f32vals(const PanelMatrix& P, const float* val32, const int* active, const double* x0, int n0,
                        const double* x1, double* partial);
float* panel_alloc_f32vals(const PanelMatrix& P);
void panel_fill_f32vals(const PanelMatrix& P, const double* csr_val, float* val32);
// bytes one panel_spmv launch moves at minimum: values + 16-bit indices,
// np (R + 1) row pointers, x once, one partial per (panel, row), 24-byte task
// descriptors
double panel_spmv_bytes(const PanelMatrix& P, int value_bytes);

// ---- hybrid operators: panels for the dense part, CSR for the rest ----------
// An operator whose nonzeros concentrate in some column blocks and rows (the
// lasso's A = [[D, -I, 0], [I, 0, -I], [I, 0, I]]: D's 20000 columns hold
// 94 % of the entries; its K = [P A'] likewise in its first 20000 rows) is
// split into
//   * panels over the rows [0, Rp) and the column blocks holding at least one
//     entry per such row on average (Rp: where the trailing run of rows with
//     at most 4 entries begins), and
//   * a remainder CSR of every other entry (original columns), which the
//     combine kernel adds to the row's panel partials.
// Used where plain panels do not pay (panel_worth) and the split moves at
// least half the entries into panels and is estimated 20 % cheaper than the
// row-block CSR kernel (streams at ~5 TB/s, random gathers at ~1.5 TB/s).
struct HybridOp {
  PanelMatrix pan;  // rows [0, Rp), selected column blocks
  int R, C, Rp;
  long long rnnz;   // remainder entries
  int* rrp;         // R + 1 remainder row pointers
  int* rci;         // rnnz: original column
  int* rsrc;        // rnnz: CSR slot the value comes from
  float* rv32;      // remainder values (FP32; also the FP32 copy of an FP64 hybrid)
  double* rv64;     // remainder values (FP64 hybrids)
  float* dc32;      // FP32 copy of an FP64 hybrid's panel values (defect correction), optional
  long long ver;
};
// 0: built; 1: does not pay (H zeroed); -1: allocation failure.
int hybrid_build(HybridOp* H, int R, int C, int bits, const int* rp, const int* ci);
// An FP64 hybrid also keeps FP32 copies of its values (panel and remainder).
int hybrid_add_f32(HybridOp* H);
void hybrid_free(HybridOp* H);
// values from the source CSR's FP64 value array (device), on the backend stream
void hybrid_refresh_values(HybridOp* H, const double* csr_val);

// FP64 panels of a general (non-triu) matrix M (transpose = 0: m x n) or of
// its transpose (transpose = 1: n x m), owned by M: built on the first call
// when they pay (panel_worth), refreshed to M's current values on every call.
// Null when they do not pay, for triu matrices, or when the build failed.
const PanelMatrix* matrix_panels64(const OSQPMatrix* M, int transpose);

// ---- combining the per-panel partials of a row -------------------------------
// A group of G consecutive lanes serves one row: lane j adds the partials of
// panels [j np / G, (j + 1) np / G) in panel order (loads four at a time), and
// a fixed xor butterfly adds the G sums (every lane of the group ends with the
// same bits).  Lane 0 of the group runs the epilogue.  Deterministic.
template <typename TX, int G>
__device__ __forceinline__ TX panel_sum(const TX* __restrict__ partial, int np, int R, int row, int j) {
  const int p0 = (np * j) / G, p1 = (np * (j + 1)) / G;
  TX s = 0;
  int p = p0;
  for (; p + 4 <= p1; p += 4) {
    const TX a0 = partial[(long long)p * R + row], a1 = partial[(long long)(p + 1) * R + row];
    const TX a2 = partial[(long long)(p + 2) * R + row], a3 = partial[(long long)(p + 3) * R + row];
    s += a0;
    s += a1;
    s += a2;
    s += a3;
  }
  for (; p < p1; ++p) s += partial[(long long)p * R + row];
#pragma unroll
  for (int off = 1; off < G; off <<= 1) s += __shfl_xor(s, off, G);
  return s;
}

// lanes per row: enough for about 256K threads, at least 4 panels per lane
inline int comb_group(int np, int rows) {
  int G = 1;
  while (G < 8 && (long long)rows * G < 262144 && np >= 4 * G) G *= 2;
  return G;
}

#define COMB_ROWS(i, j, R, G)                                                               \
  for (long long _t = (long long)blockIdx.x * blockDim.x + threadIdx.x, _stride = (long long)gridDim.x * blockDim.x; \
       _t < (long long)(R) * (G); _t += _stride)                                           \
    if (const int i = (int)(_t / (G)), j = (int)(_t % (G)); true)


}  // namespace osqp_hip
