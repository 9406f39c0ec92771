// hip_internal.h — shared definitions of the algebra/hip backend (gfx950).
//
// Storage model: every OSQP vector is an FP64 array in HBM; matrices keep
// CSR copies for the operator and its transpose (see matrix.hip).  All work
// is queued on one backend stream per process; calls returning host scalars
// synchronise that stream, nothing else does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../../include/osqp_amd_algebra.h"

namespace osqp_hip {

constexpr int kBlock = 256;          // threads per block for streaming kernels
constexpr int kMaxReduceBlocks = 2048;  // cap on per-block partials of a reduction
constexpr int kWave = 64;            // CDNA wavefront width

struct Runtime {
  int device = -1;
  int refcount = 0;
  hipStream_t stream = nullptr;
  double* h_scalars = nullptr;     // host-coherent slots for reduction results
  double* dh_scalars = nullptr;    // their device address
  double* d_partials = nullptr;    // per-block partials (4 * kMaxReduceBlocks doubles)
  unsigned* d_ticket = nullptr;    // finished-block counter of the running reduction
  double red_seq = 0.0;            // sequence number of the last reduction
  int last_error = 0;
  int num_cu = 0;
  char device_name[256] = {0};
};

Runtime& rt();

// Records (and prints once) a HIP failure; the public API keeps running and
// reports errors through return codes, never exit().
int check(hipError_t e, const char* what, const char* file, int line);
#define HIPCHK(x) ::osqp_hip::check((x), #x, __FILE__, __LINE__)
#define LAUNCHCHK() ::osqp_hip::check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

inline int grid_for(long long n, int per_block = kBlock, int cap = 65535 * 8) {
  long long g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

inline int reduce_grid(long long n) {
  long long g = (n + kBlock * 4 - 1) / (kBlock * 4);
  if (g < 1) g = 1;
  if (g > kMaxReduceBlocks) g = kMaxReduceBlocks;
  return (int)g;
}

void* dev_alloc(size_t bytes);           // stream-ordered device allocation
void dev_free(void* p);
void upload(void* dst, const void* src, size_t bytes);    // host or device src
void download(void* dst, const void* src, size_t bytes);  // host or device dst; syncs

// ---- block / wavefront reductions ----------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, kWave);
    v = o > v ? o : v;
  }
  return v;
}

// Block totals: every thread combines the wavefront totals itself, in
// wavefront order, so all threads get the same bits without a broadcast
// round; the final barrier lets the caller reuse lds at once.
template <typename T>
__device__ inline T block_sum(T v, T* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  T r = 0;
  for (int i = 0; i < (int)(blockDim.x / kWave); ++i) r += lds[i];
  __syncthreads();
  return r;
}

template <typename T>
__device__ inline T block_max(T v, T* lds) {
  v = wave_max(v);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  T r = lds[0];
  for (int i = 1; i < (int)(blockDim.x / kWave); ++i) r = lds[i] > r ? lds[i] : r;
  __syncthreads();
  return r;
}

// block_sum(s) and block_max(m) in one round (lds holds 2 * blockDim.x / 64
// values); the same bits as the two calls.
__device__ inline void block_sum_max(double& s, double& m, double* lds) {
  s = wave_sum(s);
  m = wave_max(m);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave, nw = (int)(blockDim.x / kWave);
  if (lane == 0) {
    lds[w] = s;
    lds[nw + w] = m;
  }
  __syncthreads();
  double rs = 0, rm = lds[nw];
  for (int i = 0; i < nw; ++i) rs += lds[i];
  for (int i = 1; i < nw; ++i) rm = lds[nw + i] > rm ? lds[nw + i] : rm;
  __syncthreads();
  s = rs;
  m = rm;
}

// Sum of a group of G consecutive lanes (G a power of two <= 64); every lane
// of the group receives the total.
template <int G, typename T>
__device__ inline T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

template <int G, typename T>
__device__ inline T group_max(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, G);
    v = o > v ? o : v;
  }
  return v;
}

// Unroll depth of the CSR row loops: each lane keeps kUnroll independent
// (index, value) loads in flight before it gathers.
constexpr int kUnroll = 4;

// Chooses lanes-per-row G for a CSR row kernel from the mean row length, so
// that one unrolled trip (G * kUnroll elements) covers a typical row.
inline int lanes_for_rows(long long nnz, long long rows) {
  double avg = rows > 0 ? (double)nnz / (double)rows : 0.0;
  double per_lane_trip = avg / kUnroll;
  if (per_lane_trip <= 4) return 4;
  if (per_lane_trip <= 8) return 8;
  if (per_lane_trip <= 16) return 16;
  if (per_lane_trip <= 32) return 32;
  return 64;
}

// Streaming loads of CSR index/value arrays: each element is read once per
// SpMV, so they bypass cache retention (keeping the gathered vector in L2).
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Partial dot product of one CSR row [beg, end) for one lane of a G-lane
// group: kUnroll predicated loads of (column, value) are issued before the
// gathers so that every lane has kUnroll memory requests in flight.  Lanes
// past the row end load nothing and gather element 0 with weight 0.
template <int G, typename TV, typename TX, typename Gather>
__device__ __forceinline__ TX csr_row_dot(const int* __restrict__ ci, const TV* __restrict__ val, int beg, int end,
                                         int lane, Gather gather) {
  TX acc = 0;
  for (int k = beg + lane; k < end; k += kUnroll * G) {
    int c[kUnroll];
    TV v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int idx = k + u * G;
      const bool ok = idx < end;
      c[u] = ok ? ld_stream(ci + idx) : 0;
      v[u] = ok ? ld_stream(val + idx) : (TV)0;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) acc += (TX)v[u] * gather(c[u]);
  }
  return acc;
}

// Grid of a G-lanes-per-row kernel: one group per row up to `cap` blocks,
// then the kernel strides over rows.
inline int rows_grid(long long rows, int G, int cap) {
  const int rpb = kBlock / G;
  long long g = (rows + rpb - 1) / rpb;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

constexpr int kRowGridCap = 2048;  // 8192 wavefronts: every SIMD of 256 CUs 8-deep

struct PanelMatrix;

}  // namespace osqp_hip

// ---- concrete types behind the opaque handles -----------------------------
struct OSQPVectorf_ {
  double* d;
  OSQPInt length;
  int is_view;
};

struct OSQPVectori_ {
  OSQPInt* d;
  OSQPInt length;
};

// A sparse matrix of the backend.  For a general matrix A (m x n):
//   rp/ci/val   : CSR of A (m rows)            — operator A x
//   trp/tci/tval: CSR of A' (n rows) = the input CSC arrays  — operator A' y
// For the cost matrix P given by its upper triangle (is_triu):
//   rp/ci/val   : CSR of the full symmetric P (n rows), rows sorted by column
//   trp/tci/tval: unused (P' = P)
// csc2csr[k] / csc2mirror[k] map entry k of the input CSC onto its slot(s) in
// the CSR copy (mirror = transposed slot of an off-diagonal triu entry, -1 on
// the diagonal); csc2t[k] = k for A (the transpose *is* the CSC).
// val32/tval32 are FP32 shadows used by the FP32 PCG; they are regenerated
// from the FP64 values whenever `fp32_dirty` is set.
struct OSQPMatrix_ {
  OSQPInt m, n;
  int is_triu;
  OSQPInt nnz_in;    // nnz of the input CSC (triu part for P)
  OSQPInt nnz;       // nnz of the CSR copy (full for P)
  int *rp, *ci;
  double* val;
  int *trp, *tci;
  double* tval;
  int *csc2csr, *csc2mirror;
  int* diag_pos;     // P only: slot of (i,i) in the full CSR, -1 if absent
  float *val32, *tval32;
  int fp32_dirty;
  int lanes, tlanes; // lanes per row for A and A' kernels
  long long version; // bumped on every value change
  // row-block partitions (nonzero-balanced) of the CSR copy and of A'
  int rb_n, trb_n;
  int *rb_rows, *trb_rows;
  // FP64 column panels of A and of A' (general matrices; matrix_panels64)
  struct osqp_hip::PanelMatrix* pan64[2];
  int pan_state[2];   // 0: not tried, 1: built, -1: not used
  double* pan_part;   // per-(panel, row) partials of the panel products
  size_t pan_part_len;
};
