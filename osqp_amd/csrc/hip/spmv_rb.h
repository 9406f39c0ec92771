// spmv_rb.h — row-block CSR SpMV for gfx950 (CSR-stream with nonzero-balanced
// row blocks).
//
// Rows are partitioned on the host into blocks holding at most kCap
// nonzeros (a row longer than kCap forms a block of its own).  A 256-thread
// workgroup takes one block at a time (grid-stride over blocks with a capped
// grid):
//   1. stream the block's nonzero range [k0, k1) through the window
//      [k0 & ~3, (k0 & ~3) + kChunk): each lane loads 4 consecutive (index,
//      value) pairs with one 16-byte index load and one or two 16-byte value
//      loads, twice (the address unit's cost is per instruction, so wide
//      loads move more per slot); kCap = kChunk - 3 leaves room for the
//      window's alignment slack.  Gather x and park the products in LDS at
//      their window offset;
//   2. reduce each row's products out of LDS — one thread per row for short
//      rows, otherwise a group of L lanes per row, L the largest power of two
//      (at most 64) with L * rows <= 256 — in a fixed order, so results are
//      deterministic and independent of the grid size;
//   3. hand each row total to an epilogue functor (scaling, fused vector
//      terms, block partials of dot products / norms).
// A block made of one long row is reduced by the whole workgroup over
// repeated chunk passes.  This is the stream variant of Greathouse & Daga,
// "Efficient sparse matrix-vector multiplication on GPUs using the CSR storage
// format" (SC 2014), with balancing by nonzero count.
#pragma once

#include "hip_internal.h"

namespace osqp_hip {

constexpr int kChunk = 2048;                   // LDS product slots (window length)
constexpr int kRbVec = 4;                      // pairs per lane per wide load
constexpr int kCap = kChunk - (kRbVec - 1);    // nonzeros per short row block
constexpr int kPerThread = kChunk / kBlock;
constexpr int kRbLoads = kChunk / (kBlock * kRbVec);

typedef int rb_i32x4 __attribute__((ext_vector_type(4)));
typedef float rb_f32x4 __attribute__((ext_vector_type(4)));
typedef double rb_f64x2 __attribute__((ext_vector_type(2)));

// Row-block partition of a CSR matrix (host builder in matrix.hip).
// Host builder of the partition: boundaries b_0 = 0 < b_1 < ... < b_nb = rows
// such that every block [b_i, b_{i+1}) holds at most kCap nonzeros and at
// most kChunk rows, or is a single row longer than kCap.  Integer-exact.
int build_row_blocks(const int* rp, int rows, int* out /* capacity rows + 1 */);

// Block-level driver.  TI: column index type (int, or uint16_t for panel-local
// indices).  Gather(c) -> TX value of the multiplied vector at column c;
// Epi(row, sum) consumes the row total (called by exactly one thread per row).
// prod: kChunk LDS slots, 16-byte aligned.
template <typename TV, typename TX, typename TI, typename Gather, typename Epi>
__device__ __forceinline__ void rb_spmv_block(int rb, int re, const int* __restrict__ rp,
                                              const TI* __restrict__ ci, const TV* __restrict__ val,
                                              TX* __restrict__ prod, Gather gather, Epi epi) {
  const int tid = threadIdx.x;
  const int k0 = rp[rb];
  const int k1 = rp[re];
  const int len = k1 - k0;
  if (re - rb == 1 && len > kCap) {
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
  // 1. stream the window: issue all index/value loads, then gather.  Pairs
  // in [base, k0) belong to the previous block (in bounds) and are masked;
  // a lane whose 4 pairs reach past k1 loads them one by one.
  const int base = k0 & ~(kRbVec - 1);
  constexpr bool kWide = sizeof(TI) == 4 && (sizeof(TV) == 4 || sizeof(TV) == 8);
  const bool wide = kWide && ((((size_t)ci) | ((size_t)val)) & 15) == 0;
  if (wide) {
    int c[kRbLoads][kRbVec];
    TV v[kRbLoads][kRbVec];
#pragma unroll
    for (int j = 0; j < kRbLoads; ++j) {
      const int q = base + (j * kBlock + tid) * kRbVec;
      if (q + kRbVec <= k1) {
        const rb_i32x4 cv = ld_stream((const rb_i32x4*)(const void*)(ci + q));
#pragma unroll
        for (int u = 0; u < kRbVec; ++u) c[j][u] = cv[u];
        if constexpr (sizeof(TV) == 4) {
          const rb_f32x4 vv = ld_stream((const rb_f32x4*)(const void*)(val + q));
#pragma unroll
          for (int u = 0; u < kRbVec; ++u) v[j][u] = (TV)vv[u];
        } else {
          const rb_f64x2 va = ld_stream((const rb_f64x2*)(const void*)(val + q));
          const rb_f64x2 vb = ld_stream((const rb_f64x2*)(const void*)(val + q + 2));
          v[j][0] = (TV)va[0];
          v[j][1] = (TV)va[1];
          v[j][2] = (TV)vb[0];
          v[j][3] = (TV)vb[1];
        }
      } else {
#pragma unroll
        for (int u = 0; u < kRbVec; ++u) {
          const bool ok = q + u < k1;
          c[j][u] = ok ? (int)ld_stream(ci + q + u) : 0;
          v[j][u] = ok ? ld_stream(val + q + u) : (TV)0;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kRbLoads; ++j) {
      const int q = base + (j * kBlock + tid) * kRbVec;
      TX pr[kRbVec];
#pragma unroll
      for (int u = 0; u < kRbVec; ++u)
        pr[u] = (q + u >= k0 && q + u < k1) ? (TX)v[j][u] * gather(c[j][u]) : (TX)0;
      TX* dst = prod + (q - base);
      if constexpr (sizeof(TX) == 4) {
        *(rb_f32x4*)(void*)dst = rb_f32x4{(float)pr[0], (float)pr[1], (float)pr[2], (float)pr[3]};
      } else {
        *(rb_f64x2*)(void*)dst = rb_f64x2{(double)pr[0], (double)pr[1]};
        *(rb_f64x2*)(void*)(dst + 2) = rb_f64x2{(double)pr[2], (double)pr[3]};
      }
    }
  } else {
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
      if (e < len) prod[k0 - base + e] = (TX)v[j] * gather(c[j]);
    }
  }
  __syncthreads();
  // 2. reduce rows out of LDS (slot of entry k: k - base)
  const int nrows = re - rb;
  if (len <= 16 * nrows) {
    // short rows: one thread per row, sequential in column order
    for (int r = tid; r < nrows; r += kBlock) {
      const int s = rp[rb + r] - base, e = rp[rb + r + 1] - base;
      TX acc = 0;
      for (int k = s; k < e; ++k) acc += prod[k];
      epi(rb + r, acc);
    }
  } else {
    // longer rows: L lanes per row (here rows < 128, so L >= 2); the lanes
    // of a group share their row, so the group's shuffles see only lanes of
    // the same group, all active together
    int g = kBlock / nrows;
    g = g < 1 ? 1 : g;
    const int lgL = g >= kWave ? 6 : 31 - __builtin_clz(g);
    const int L = 1 << lgL;
    const int lane = tid & (L - 1), grp = tid >> lgL, ngrp = kBlock >> lgL;
    for (int r = grp; r < nrows; r += ngrp) {
      const int s = rp[rb + r] - base, e = rp[rb + r + 1] - base;
      TX acc = 0;
      for (int k = s + lane; k < e; k += L) acc += prod[k];
      for (int o = L >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
      if (lane == 0) epi(rb + r, acc);
    }
  }
  __syncthreads();
}

}  // namespace osqp_hip
