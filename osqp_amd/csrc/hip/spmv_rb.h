// spmv_rb.h — row-block CSR SpMV for gfx950 (CSR-stream with nonzero-balanced
// row blocks).
//
// Rows are partitioned on the host into blocks holding at most kChunk
// nonzeros (a row longer than kChunk forms a block of its own).  A 256-thread
// workgroup takes one block at a time (grid-stride over blocks with a capped
// grid):
//   1. stream the block's contiguous nonzero range with fully coalesced,
//      unrolled loads (kChunk / 256 = 8 (index, value) pairs per thread in
//      flight), gather x and park the products in LDS;
//   2. reduce each row's products out of LDS — one thread per row for short
//      rows, one wavefront per row otherwise — in a fixed order, so results
//      are deterministic and independent of the grid size;
//   3. hand each row total to an epilogue functor (scaling, fused vector
//      terms, block partials of dot products / norms).
// A block made of one long row is reduced by the whole workgroup over
// repeated chunk passes.  This is the stream variant of Greathouse & Daga,
// "Efficient sparse matrix-vector multiplication on GPUs using the CSR storage
// format" (SC 2014), with balancing by nonzero count.
#pragma once

#include "hip_internal.h"

namespace osqp_hip {

constexpr int kChunk = 2048;            // nonzeros per row block
constexpr int kPerThread = kChunk / kBlock;

// Row-block partition of a CSR matrix (host builder in matrix.hip).
// Host builder of the partition: boundaries b_0 = 0 < b_1 < ... < b_nb = rows
// such that every block [b_i, b_{i+1}) holds at most kChunk nonzeros and at
// most kChunk rows, or is a single row longer than kChunk.  Integer-exact.
int build_row_blocks(const int* rp, int rows, int* out /* capacity rows + 1 */);

// Block-level driver.  TI: column index type (int, or uint16_t for panel-local
// indices).  Gather(c) -> TX value of the multiplied vector at column c;
// Epi(row, sum) consumes the row total (called by exactly one thread per row).
template <typename TV, typename TX, typename TI, typename Gather, typename Epi>
__device__ __forceinline__ void rb_spmv_block(int rb, int re, const int* __restrict__ rp,
                                              const TI* __restrict__ ci, const TV* __restrict__ val,
                                              TX* __restrict__ prod, Gather gather, Epi epi) {
  const int tid = threadIdx.x;
  const int k0 = rp[rb];
  const int k1 = rp[re];
  const int len = k1 - k0;
  if (re - rb == 1 && len > kChunk) {
    // one long row: whole-workgroup reduction over chunk passes (the LDS
    // product buffer holds the per-wavefront sums)
    TX acc = 0;
    for (int base = k0; base < k1; base += kChunk) {
#pragma unroll
      for (int j = 0; j < kPerThread; ++j) {
        const int k = base + j * kBlock + tid;
        if (k < k1) acc += (TX)ld_stream(val + k) * gather((int)ld_stream(ci + k));
      }
    }
    acc = wave_sum(acc);
    if ((tid & (kWave - 1)) == 0) prod[tid / kWave] = acc;
    __syncthreads();
    if (tid == 0) {
      TX s = 0;
      for (int w = 0; w < kBlock / kWave; ++w) s += prod[w];
      epi(rb, s);
    }
    __syncthreads();
    return;
  }
  // 1. stream the chunk: issue all index/value loads, then gather
  int c[kPerThread];
  TV v[kPerThread];
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const int e = j * kBlock + tid;
    const bool ok = e < len;
    c[j] = ok ? (int)ld_stream(ci + k0 + e) : 0;
    v[j] = ok ? ld_stream(val + k0 + e) : (TV)0;
  }
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const int e = j * kBlock + tid;
    if (e < len) prod[e] = (TX)v[j] * gather(c[j]);
  }
  __syncthreads();
  // 2. reduce rows out of LDS
  const int nrows = re - rb;
  if (len <= 16 * nrows) {
    // short rows: one thread per row, sequential in column order
    for (int r = tid; r < nrows; r += kBlock) {
      const int s = rp[rb + r] - k0, e = rp[rb + r + 1] - k0;
      TX acc = 0;
      for (int k = s; k < e; ++k) acc += prod[k];
      epi(rb + r, acc);
    }
  } else {
    // longer rows: one wavefront per row
    const int lane = tid & (kWave - 1), wv = tid / kWave;
    for (int r = wv; r < nrows; r += kBlock / kWave) {
      const int s = rp[rb + r] - k0, e = rp[rb + r + 1] - k0;
      TX acc = 0;
      for (int k = s + lane; k < e; k += kWave) acc += prod[k];
      acc = wave_sum(acc);
      if (lane == 0) epi(rb + r, acc);
    }
  }
  __syncthreads();
}

}  // namespace osqp_hip
