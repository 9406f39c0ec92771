// pcg.hip — matrix-free Jacobi-preconditioned CG on the reduced KKT system
//     K x = (P + sigma I + A' diag(rho) A) x = b1 + A' (rho o b2)
// behind OSQP's linsys plugin vtable (include/private/types.h:243-279).
//
// Reference: algebra/cuda/lin_sys/indirect/cuda_pcg_interface.cu:32-361 and
// cuda_pcg.cu:50-285; CPU statement of K in algebra/_common/reduced_kkt.c.
//
// MI355X design:
//   * Device-resident control.  The loop state (r'y, ||r||, alpha, beta, eps,
//     iteration count, active flag) lives in a device struct.  The kernel that
//     consumes a reduction finishes it itself, identically in every block
//     (alpha inside the update, beta and the stop test inside the
//     search-direction update), so a CG iteration is six launches: two panel
//     SpMVs, their two combines, the update and the loop head.  Every CG
//     kernel returns at once when `active` is cleared.  The host never waits on the GPU inside a
//     solve: it queues CG iterations ahead and reads a per-iteration flag the
//     loop-head kernel writes to host-coherent memory, so it stops queueing a
//     couple of iterations after convergence.  The reference host-syncs every
//     CG iteration (cuda_pcg.cu:188,199).
//   * The tolerance schedule (cuda_pcg_interface.cu:32-64) runs on the device,
//     including the zero-iteration counter that halves the reduction factor.
//   * Defect-correction form for both arithmetics: the iterate x (warm start),
//     the right-hand side and the residual r = K x - rhs are FP64 with the
//     FP64 matrix values (the exact K); CG runs on the correction system
//     K d = -r/||r|| and x += ||r|| d.  In exact arithmetic this is CG started
//     from x, i.e. the reference iteration.
//       FP64 ("plain"): CG values and vectors in double.
//       FP32 ("mixed"): CG on FP32-stored values (8 bytes per nonzero with the
//         int32 index), FP32 vectors and lane accumulation, Kahan-compensated
//         dot products.  Because the residual uses the exact K, neither the
//         FP32 round-off nor the FP32 rounding of the matrix limits the
//         accuracy of x: it is classic mixed-precision iterative refinement.
//   * Traffic per ADMM iteration: after the first solve the residual reuses
//     A x from the previous solve's z = A x (r = sigma x + P x - b1 +
//     A' (rho o (A x - b2))), so one solve costs (CG iterations + 1) passes
//     over A and A' instead of (CG iterations + 2).
//   * The operator is two row-block SpMVs (spmv_rb.h): w = rho o (A p) over
//     A, and K p = sigma p + [P A'] [p; w] over one combined CSR of P and A'
//     built at setup, so each CG iteration is two streaming passes.
#include <math.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <vector>

#include <type_traits>

#include "hip_internal.h"
#include "spmv_rb.h"
#include "panel.h"
#include "../profilers.h"

using namespace osqp_hip;

namespace {

inline hipStream_t S() { return rt().stream; }

#define GRID_STRIDE(i, n) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

struct PcgState {
  double eps;          // tolerance of the current solve (absolute, on K x = rhs)
  double eps_cmp;      // tolerance compared with the CG residual (eps / scale)
  double scale;        // defect-correction scale ||r0||_inf
  double eps_prev;     // tolerance schedule state
  double red_factor;   // tolerance schedule state
  double rhs_norm;
  double rty[2];       // r'y of the iterate after j CG iterations, in slot j & 1
  double rnorm;        // ||r||_inf of the CG residual
  double pKp, alpha, beta;
  int iter;            // CG iterations of the current solve
  int active;          // 1 while the CG loop runs
  int done;            // 1 once the current solve's CG loop has stopped
  int zero_iters;      // consecutive solves with 0 CG iterations
  int solves;          // solves performed
  long long total_iters;
};

// Kahan-compensated FP32 accumulator (float) / plain FP64 (double)
template <typename V>
struct Acc;
template <>
struct Acc<double> {
  double s = 0.0;
  __device__ void add(double v) { s += v; }
  __device__ double value() const { return s; }
};
template <>
struct Acc<float> {
  float s = 0.0f, c = 0.0f;
  __device__ void add(float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ double value() const { return (double)s - (double)c; }
};

__device__ inline double rho_at(const double* rho, double rho_sc, int i) { return rho ? rho[i] : rho_sc; }

// ---- row-block (nonzero-balanced) versions of the operator kernels -----------
// w = rho o (A v)
template <typename TV, typename TX>
__global__ __launch_bounds__(kBlock) void k_rb_Av_rho(const PcgState* st, const int* __restrict__ rows, int nb,
                                                      const int* __restrict__ rp, const int* __restrict__ ci,
                                                      const TV* __restrict__ val, const TX* __restrict__ v,
                                                      const double* __restrict__ rho, double rho_sc,
                                                      TX* __restrict__ w) {
  if (!st->active) return;
  __shared__ __attribute__((aligned(16))) TX prod[kChunk];
  for (int b = blockIdx.x; b < nb; b += gridDim.x)
    rb_spmv_block<TV, TX>(rows[b], rows[b + 1], rp, ci, val, prod, [&](int c) { return v[c]; },
                          [&](int row, TX sum) { w[row] = (TX)rho_at(rho, rho_sc, row) * sum; });
}

// Kv = sigma v + [P A'] [v; w] (- sub) over the combined K-operator CSR.
// sub == nullptr: partials of v'Kv; otherwise partials of max|Kv - sub|.
template <typename TV, typename TX>
__global__ __launch_bounds__(kBlock) void k_rb_Kv(const PcgState* st, const int* __restrict__ rows, int nb,
                                                  const int* __restrict__ krp, const int* __restrict__ kci,
                                                  const TV* __restrict__ kval, int n, double sigma,
                                                  const TX* __restrict__ v, const TX* __restrict__ w,
                                                  const double* __restrict__ sub, TX* __restrict__ Kv,
                                                  double* partials) {
  if (!sub && !st->active) return;  // the residual (sub given) runs outside the CG loop
  __shared__ __attribute__((aligned(16))) TX prod[kChunk];
  __shared__ double lds[kBlock / kWave];
  const TX sg = (TX)sigma;
  Acc<TX> dot;
  double mx = 0.0;
  for (int b = blockIdx.x; b < nb; b += gridDim.x)
    rb_spmv_block<TV, TX>(rows[b], rows[b + 1], krp, kci, kval, prod,
                          [&](int c) { return c < n ? v[c] : w[c - n]; },
                          [&](int row, TX sum) {
                            TX vi = v[row];
                            TX kv = sg * vi + sum;
                            if (sub) {
                              kv -= (TX)sub[row];
                              double a = fabs((double)kv);
                              mx = a > mx ? a : mx;
                            } else {
                              dot.add(vi * kv);
                            }
                            Kv[row] = kv;
                          });
  double t = sub ? block_max(mx, lds) : block_sum(dot.value(), lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// rhs = b1 + A' (rho o b2) over the K-operator CSR (P columns weigh zero)
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_rb_rhs(const int* __restrict__ rows, int nb, const int* __restrict__ krp,
                                                   const int* __restrict__ kci, const TV* __restrict__ kval, int n,
                                                   const double* __restrict__ b, const double* __restrict__ rho,
                                                   double rho_sc, double* __restrict__ rhs, double* partials) {
  __shared__ __attribute__((aligned(16))) double prod[kChunk];
  __shared__ double lds[kBlock / kWave];
  const double* b2 = b + n;
  double mx = 0.0;
  for (int blk = blockIdx.x; blk < nb; blk += gridDim.x)
    rb_spmv_block<TV, double>(rows[blk], rows[blk + 1], krp, kci, kval, prod,
                              [&](int c) { return c < n ? 0.0 : rho_at(rho, rho_sc, c - n) * b2[c - n]; },
                              [&](int row, double sum) {
                                double r = b[row] + sum;
                                rhs[row] = r;
                                double a = fabs(r);
                                mx = a > mx ? a : mx;
                              });
  double t = block_max(mx, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// b[n:n+m] = A x (FP64 accumulation); ax = A x; polishing: rho_pol (A x - b2)
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_rb_finish_z(const int* __restrict__ rows, int nb,
                                                        const int* __restrict__ rp, const int* __restrict__ ci,
                                                        const TV* __restrict__ val, const double* __restrict__ x,
                                                        double* __restrict__ b, double* __restrict__ ax, int n,
                                                        int polishing, double rho_pol) {
  __shared__ __attribute__((aligned(16))) double prod[kChunk];
  for (int blk = blockIdx.x; blk < nb; blk += gridDim.x)
    rb_spmv_block<TV, double>(rows[blk], rows[blk + 1], rp, ci, val, prod, [&](int c) { return x[c]; },
                              [&](int row, double sum) {
                                if (ax) ax[row] = sum;
                                b[n + row] = polishing ? (sum - b[n + row]) * rho_pol : sum;
                              });
}

// combined K-operator CSR [P A'] (n rows; A' columns shifted by n)
__global__ void k_krp(int n, const int* prp, const int* trp, int* krp) {
  GRID_STRIDE(r, n + 1) krp[r] = prp[r] + (trp ? trp[r] : 0);
}

template <typename TV>
__global__ void k_kfill(int n, const int* prp, const int* pci, const double* pval, const int* trp, const int* tci,
                        const double* tval, const int* krp, int* kci, TV* kval, int cols_too) {
  const int lane = threadIdx.x & (kWave - 1);
  for (long long r = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave; r < n;
       r += (long long)gridDim.x * blockDim.x / kWave) {
    const int o = krp[r], p0 = prp[r], p1 = prp[r + 1];
    for (int k = p0 + lane; k < p1; k += kWave) {
      kval[o + k - p0] = (TV)pval[k];
      if (cols_too) kci[o + k - p0] = pci[k];
    }
    if (trp) {
      const int o2 = o + (p1 - p0), t0 = trp[r], t1 = trp[r + 1];
      for (int k = t0 + lane; k < t1; k += kWave) {
        kval[o2 + k - t0] = (TV)tval[k];
        if (cols_too) kci[o2 + k - t0] = tci[k] + n;
      }
    }
  }
}

// w = rho o (A v)
// (ap, when given, receives A v itself: the CG's products for the running
// A d of the incremental z = A x64)
template <typename TX, int G>
__global__ void k_comb_av(const PcgState* st, int m, int np, const TX* partial, const double* rho, double rho_sc,
                          TX* w, TX* ap) {
  if (st && !st->active) return;
  COMB_ROWS(i, j, m, G) {
    const TX s = panel_sum<TX, G>(partial, np, m, i, j);
    if (j == 0) {
      w[i] = (TX)rho_at(rho, rho_sc, i) * s;
      if (ap) ap[i] = s;
    }
  }
}

// Kv = sigma v + K-panels (- sub); partials of v'Kv or of max|Kv - sub|
template <typename TX, int G>
__global__ void k_comb_kv(const PcgState* st, int n, int np, const TX* partial, double sigma, const TX* v,
                          const double* sub, TX* Kv, double* partials) {
  if (!sub && !st->active) return;  // the residual (sub given) runs outside the CG loop
  __shared__ double lds[kBlock / kWave];
  const TX sg = (TX)sigma;
  Acc<TX> dot;
  double mx = 0.0;
  COMB_ROWS(i, j, n, G) {
    const TX s = panel_sum<TX, G>(partial, np, n, i, j);
    if (j == 0) {
      const TX vi = v[i];
      TX kv = sg * vi + s;
      if (sub) {
        kv -= (TX)sub[i];
        const double a = fabs((double)kv);
        mx = a > mx ? a : mx;
      } else {
        dot.add(vi * kv);
      }
      Kv[i] = kv;
    }
  }
  const double t = sub ? block_max(mx, lds) : block_sum(dot.value(), lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// The same for n divisible by RB = 16 / sizeof(TX): a group of G lanes serves
// RB adjacent rows, and a lane loads their partials of a panel as one 16-byte
// vector (two or four times the data per load instruction of the form above;
// the address unit's cost is per instruction).  Each row's partials are added
// in the same order as by panel_sum, so Kv is bit-identical to k_comb_kv's.
template <typename TX, int G>
__global__ void k_comb_kv_rows(const PcgState* st, int n, int np, const TX* partial, double sigma, const TX* v,
                               const double* sub, TX* Kv, double* partials) {
  if (!sub && !st->active) return;  // the residual (sub given) runs outside the CG loop
  constexpr int RB = 16 / (int)sizeof(TX);
  typedef TX VT __attribute__((ext_vector_type(RB)));
  __shared__ double lds[kBlock / kWave];
  const TX sg = (TX)sigma;
  Acc<TX> dot;
  double mx = 0.0;
  const int nq = n / RB;
  COMB_ROWS(q, j, nq, G) {
    const int p0 = (np * j) / G, p1 = (np * (j + 1)) / G;
    const VT* __restrict__ col = (const VT*)(partial + (long long)q * RB);
    TX s[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) s[r] = 0;
    int p = p0;
    for (; p + 4 <= p1; p += 4) {
      const VT a0 = col[(long long)p * nq], a1 = col[(long long)(p + 1) * nq];
      const VT a2 = col[(long long)(p + 2) * nq], a3 = col[(long long)(p + 3) * nq];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        s[r] += a0[r];
        s[r] += a1[r];
        s[r] += a2[r];
        s[r] += a3[r];
      }
    }
    for (; p < p1; ++p) {
      const VT a = col[(long long)p * nq];
#pragma unroll
      for (int r = 0; r < RB; ++r) s[r] += a[r];
    }
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int off = 1; off < G; off <<= 1) s[r] += __shfl_xor(s[r], off, G);
    if (j == 0) {
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int i = q * RB + r;
        const TX vi = v[i];
        TX kv = sg * vi + s[r];
        if (sub) {
          kv -= (TX)sub[i];
          const double a = fabs((double)kv);
          mx = a > mx ? a : mx;
        } else {
          dot.add(vi * kv);
        }
        Kv[i] = kv;
      }
    }
  }
  const double t = sub ? block_max(mx, lds) : block_sum(dot.value(), lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// z = A x (FP64): b[n:n+m] (polishing: rho_pol (A x - b2)) and the A x cache;
// skipped while *when == 0 (when given)
template <int G>
__global__ void k_comb_z(const int* when, int m, int n, int np, const double* partial, double* b, double* ax,
                         int polishing, double rho_pol) {
  if (when && !*when) return;
  COMB_ROWS(i, j, m, G) {
    const double s = panel_sum<double, G>(partial, np, m, i, j);
    if (j == 0) {
      if (ax) ax[i] = s;
      b[n + i] = polishing ? (s - b[n + i]) * rho_pol : s;
    }
  }
}

// ---- hybrid operators: panel partials of rows < Rp plus the remainder CSR ------
// A group of G lanes serves a row: lane j adds its share of the row's panel
// partials (as panel_sum) and the remainder entries j, j + G, ... in order;
// a fixed butterfly adds the G sums.  Deterministic.
template <typename TR>
struct Rem {
  int Rp;          // rows with panel partials (stride of the partials)
  const int* rp;   // remainder CSR (R + 1 row pointers)
  const int* ci;   // original columns
  const TR* val;
};

template <typename TX, int G, typename TR, typename Gather>
__device__ __forceinline__ TX hyb_sum(const TX* __restrict__ partial, int np, const Rem<TR>& rm, int row, int j,
                                      Gather gather) {
  TX s = 0;
  if (row < rm.Rp) {
    const int p0 = (np * j) / G, p1 = (np * (j + 1)) / G;
    for (int p = p0; p < p1; ++p) s += partial[(long long)p * rm.Rp + row];
  }
  const int k1 = rm.rp[row + 1];
  for (int k = rm.rp[row] + j; k < k1; k += G) s += (TX)rm.val[k] * gather(rm.ci[k]);
#pragma unroll
  for (int off = 1; off < G; off <<= 1) s += __shfl_xor(s, off, G);
  return s;
}

// w = rho o (A v)
template <typename TX, int G, typename TR>
__global__ void k_hyb_av(const PcgState* st, int m, int np, const TX* partial, Rem<TR> rm, const TX* v,
                         const double* rho, double rho_sc, TX* w, TX* ap) {
  if (st && !st->active) return;
  COMB_ROWS(i, j, m, G) {
    const TX s = hyb_sum<TX, G>(partial, np, rm, i, j, [&](int c) { return v[c]; });
    if (j == 0) {
      w[i] = (TX)rho_at(rho, rho_sc, i) * s;
      if (ap) ap[i] = s;
    }
  }
}

// Kv = sigma v + [P A'] [v; w] (- sub); partials of v'Kv or of max|Kv - sub|
template <typename TX, int G, typename TR>
__global__ void k_hyb_kv(const PcgState* st, int n, int np, const TX* partial, Rem<TR> rm, double sigma, const TX* v,
                         const TX* w, const double* sub, TX* Kv, double* partials) {
  if (!sub && !st->active) return;  // the residual (sub given) runs outside the CG loop
  __shared__ double lds[kBlock / kWave];
  const TX sg = (TX)sigma;
  Acc<TX> dot;
  double mx = 0.0;
  COMB_ROWS(i, j, n, G) {
    const TX s = hyb_sum<TX, G>(partial, np, rm, i, j, [&](int c) { return c < n ? v[c] : w[c - n]; });
    if (j == 0) {
      const TX vi = v[i];
      TX kv = sg * vi + s;
      if (sub) {
        kv -= (TX)sub[i];
        const double a = fabs((double)kv);
        mx = a > mx ? a : mx;
      } else {
        dot.add(vi * kv);
      }
      Kv[i] = kv;
    }
  }
  const double t = sub ? block_max(mx, lds) : block_sum(dot.value(), lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// z = A x (FP64 sums): b[n:n+m] (polishing: rho_pol (A x - b2)) and the A x
// cache; skipped while *when == 0 (when given)
template <int G, typename TR>
__global__ void k_hyb_z(const int* when, int m, int n, int np, const double* partial, Rem<TR> rm, const double* x,
                        double* b, double* ax, int polishing, double rho_pol) {
  if (when && !*when) return;
  COMB_ROWS(i, j, m, G) {
    const double s = hyb_sum<double, G>(partial, np, rm, i, j, [&](int c) { return x[c]; });
    if (j == 0) {
      if (ax) ax[i] = s;
      b[n + i] = polishing ? (s - b[n + i]) * rho_pol : s;
    }
  }
}

// ---- tolerance (cuda_pcg_interface.cu:32-64) + loop reset ---------------------
// One thread.  rhs_norm matters only at the first ADMM iteration and in
// polishing.
__device__ void set_tolerance(PcgState* st, double rhs_norm, int admm_iter, double spr, double sdr, int polishing,
                              double tol_fraction, int reduction_threshold) {
  // zero-iteration bookkeeping of the previous solve (cuda_pcg_interface.cu:266-268)
  if (st->solves > 0) {
    if (st->iter == 0) st->zero_iters++;
    else st->zero_iters = 0;
  }
  double eps;
  if (polishing) {
    eps = fmax(rhs_norm * 1e-5 /* OSQP_CG_POLISH_TOL */, 1e-7 /* OSQP_CG_TOL_MIN */);
  } else if (admm_iter == 1) {
    st->red_factor = tol_fraction;
    st->eps_prev = (rhs_norm < 1e-7) ? 1.0 : rhs_norm * st->red_factor;
    eps = st->eps_prev;
  } else {
    if (st->zero_iters >= reduction_threshold) {
      st->red_factor /= 2;
      st->zero_iters = 0;
    }
    eps = st->red_factor * sqrt(spr * sdr);
    eps = eps < st->eps_prev ? eps : st->eps_prev;
    eps = eps > 1e-7 ? eps : 1e-7;
    st->eps_prev = eps;
  }
  st->rhs_norm = rhs_norm;
  st->eps = eps;
  st->eps_cmp = eps;
  st->scale = 1.0;
  st->iter = 0;
  st->active = 1;
  st->done = 0;
  st->solves++;
}

// Every block of the kernels below that consume a reduction forms its total
// itself, from the producer's per-block partials, in one fixed order (thread i
// takes partials i, i + 256, ...; then the block tree), so all blocks agree to
// the bit and no one-block kernel sits between producer and consumer.  Block 0
// alone records the scalars in the state.
__device__ inline double parts_max(const double* part, int nparts, double* lds) {
  double mx = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) mx = part[i] > mx ? part[i] : mx;
  return block_max(mx, lds);
}
// parts_sum(a) and parts_max(b) in one pass and one block round (lds: 2 *
// blockDim.x / 64 values); the same bits as the two calls
__device__ inline void parts_sum_max(const double* a, const double* b, int nparts, double* lds, double* sum,
                                     double* mx) {
  double s = 0.0, m = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    s += a[i];
    m = b[i] > m ? b[i] : m;
  }
  block_sum_max(s, m, lds);
  *sum = s;
  *mx = m;
}
__device__ inline double parts_sum(const double* part, int nparts, double* lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  return block_sum(s, lds);
}

// explicit-rhs path: rhs_norm from the rhs kernel's partials; one block
__global__ void k_tolerance(PcgState* st, const double* partials, int nparts, int admm_iter, double spr, double sdr,
                            int polishing, double tol_fraction, int reduction_threshold) {
  __shared__ double lds[kBlock / kWave];
  const double rhs_norm = parts_max(partials, nparts, lds);
  if (threadIdx.x == 0)
    set_tolerance(st, rhs_norm, admm_iter, spr, sdr, polishing, tol_fraction, reduction_threshold);
}

// cached path: the tolerance (block 0) and w = rho o (A x - b2) from the
// cached A x of the previous solve (m rows)
__global__ void k_tol_w(PcgState* st, int admm_iter, double spr, double sdr, double tol_fraction,
                        int reduction_threshold, int m, const double* ax, const double* b2, const double* rho,
                        double rho_sc, double* w) {
  if (blockIdx.x == 0 && threadIdx.x == 0)
    set_tolerance(st, 0.0, admm_iter, spr, sdr, 0, tol_fraction, reduction_threshold);
  GRID_STRIDE(i, m) w[i] = rho_at(rho, rho_sc, (int)i) * (ax[i] - b2[i]);
}

// scale = ||r64||_inf from the residual kernel's partials (tolerance on the
// scaled correction system: eps / scale); r = r64 / scale; y = M^-1 r; p = -y;
// d = 0; partials r'y, max|r|.  `part_res` must not alias the outputs.
template <typename V>
__global__ void k_dc_init_cg(PcgState* st, const double* part_res, int nres, int n, const double* r64, const V* dinv,
                             V* r, V* y, V* p, V* d, double* part_rty, double* part_rn, int settol, int admm_iter,
                             double spr, double sdr, double tol_fraction, int reduction_threshold) {
  __shared__ double lds2[2 * kBlock / kWave];
  double* lds = lds2;
  const double rn0 = parts_max(part_res, nres, lds);
  const double scale = (rn0 > 0.0 && isfinite(rn0)) ? rn0 : 1.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // cached path with w64 given: the tolerance update of k_tol_w, here
    if (settol) set_tolerance(st, 0.0, admm_iter, spr, sdr, 0, tol_fraction, reduction_threshold);
    st->scale = scale;
    st->eps_cmp = st->eps / scale;
  }
  const double inv = 1.0 / scale;
  Acc<V> dot;
  double mx = 0.0;
  GRID_STRIDE(i, n) {
    V ri = (V)(r64[i] * inv);
    V yi = dinv[i] * ri;
    r[i] = ri;
    y[i] = yi;
    p[i] = -yi;
    d[i] = (V)0;
    dot.add(yi * ri);
    double a = fabs((double)ri);
    mx = a > mx ? a : mx;
  }
  double t1 = dot.value(), t2 = mx;
  block_sum_max(t1, t2, lds2);
  if (threadIdx.x == 0) {
    part_rty[blockIdx.x] = t1;
    part_rn[blockIdx.x] = t2;
  }
}

// x64 += scale * d ; b[0:n] = x64 (with when_done: only once the loop has stopped)
template <typename V>
__global__ void k_dc_apply(const PcgState* st, int when_done, int n, const V* d, double* x64, double* b) {
  if (when_done && !st->done) return;
  const double s = st->scale;
  const int ran = st->iter > 0;
  GRID_STRIDE(i, n) {
    double xi = x64[i];
    if (ran) xi += s * (double)d[i];
    x64[i] = xi;
    b[i] = xi;
  }
}

// The same, with the incremental z = A x64 of the FP32 engine in one launch:
// rows i < n as k_dc_apply; rows n + i: z = ax + scale u (u = A d summed over
// the CG iterations from their own A p products), ax = z, b[n + i] = z.
__global__ void k_dc_apply_incr(const PcgState* st, int when_done, int n, int m, const float* d, const float* u,
                                double* x64, double* ax, double* b) {
  if (when_done && !st->done) return;
  const double s = st->scale;
  const int ran = st->iter > 0;
  GRID_STRIDE(k, (long long)n + m) {
    const bool xrow = k < n;
    const long long i = xrow ? k : k - n;
    double* t = xrow ? x64 : ax;
    double v = t[i];
    if (ran) v += s * (double)(xrow ? d[i] : u[i]);
    t[i] = v;
    b[k] = v;
  }
}

// ---- CG loop control ------------------------------------------------------------
// `flags` (host-coherent, may be null): flags[j] = 2 * solve + active after the
// loop head of CG iteration j, read by the host to stop queueing iterations.
// Loop state written inside a kernel is read by that kernel's other blocks
// only where the value cannot have changed (r'y lives in two slots by
// iteration parity; the iteration count is the launch argument j).

// stop test at the head of CG iteration j (one thread)
__device__ inline int loop_go(const PcgState* st, double rty, double rn, int j, int max_iter, int guard) {
  int go = (rn > st->eps_cmp && j < max_iter) ? 1 : 0;
  // the mixed path stops on a non-finite or non-positive r'y (FP32 breakdown)
  if (guard && !(rty > 0.0 && isfinite(rty))) go = 0;
  return go;
}

// CG iteration j: alpha = r'y / p'Kp from the operator's p'Kp partials; then
// d += alpha p; r += alpha Kp; y = M^-1 r; partials r'y, max|r|.  On an FP32
// breakdown (p'Kp not positive and finite) every block skips the update and
// block 0 stops the loop.  Iteration 0 also runs the loop head of the
// defect-correction init (r'y and max|r| from its partials init_rty /
// init_rn, the stop test, flag 0), so no one-block kernel sits between the
// init and the first operator; when that head stops the loop, the operator
// of iteration 0 has run for nothing and the update is skipped.
template <typename V>
__global__ void k_update(PcgState* st, const double* part_pkp, int npkp, int j, int guard, int n, V* x, V* r, V* y,
                         const V* p, const V* Kp, const V* dinv, double* part_rty, double* part_rn,
                         const double* init_rty, const double* init_rn, int ninit, int max_iter, int* flags, int m,
                         const V* ap, float* u) {
  if (!st->active) return;
  __shared__ double lds2[2 * kBlock / kWave];
  double* lds = lds2;
  // the first grid-stride element's operands do not depend on alpha: they
  // are loaded before the partial sums, so the two memory latencies overlap
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  V x0 = 0, p0 = 0, r0 = 0, kp0 = 0, d0 = 0;
  if (i0 < n) {
    x0 = x[i0];
    p0 = p[i0];
    r0 = r[i0];
    kp0 = Kp[i0];
    d0 = dinv[i0];
  }
  double rty_j;
  if (j == 0) {
    double rty0, rn0;
    parts_sum_max(init_rty, init_rn, ninit, lds2, &rty0, &rn0);
    const int go = loop_go(st, rty0, rn0, 0, max_iter, guard);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->rty[0] = rty0;
      st->rnorm = rn0;
      if (!go) {
        st->active = 0;
        st->done = 1;
      }
      if (flags) flags[0] = 2 * st->solves + go;
    }
    if (!go) return;
    rty_j = rty0;
  } else {
    rty_j = st->rty[j & 1];
  }
  const double pkp = parts_sum(part_pkp, npkp, lds);
  const double alpha64 = rty_j / pkp;
  const bool broke = guard && !(pkp > 0.0 && isfinite(pkp));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pKp = pkp;
    st->alpha = alpha64;
    if (broke) {
      st->active = 0;
      st->done = 1;
    }
  }
  if (broke) return;
  const V alpha = (V)alpha64;
  Acc<V> dot;
  double mx = 0.0;
  if (i0 < n) {
    x[i0] = x0 + alpha * p0;
    V ri = r0 + alpha * kp0;
    V yi = d0 * ri;
    r[i0] = ri;
    y[i0] = yi;
    dot.add(yi * ri);
    double a = fabs((double)ri);
    mx = a > mx ? a : mx;
  }
  for (long long i = i0 + stride; i < n; i += stride) {
    x[i] += alpha * p[i];
    V ri = r[i] + alpha * Kp[i];
    V yi = dinv[i] * ri;
    r[i] = ri;
    y[i] = yi;
    dot.add(yi * ri);
    double a = fabs((double)ri);
    mx = a > mx ? a : mx;
  }
  // u += alpha A p, the running A d (same rounded alpha as d's update;
  // iteration 0 assigns it)
  if (u) {
    if (j == 0)
      for (long long i = i0; i < m; i += stride) u[i] = (float)alpha * (float)ap[i];
    else
      for (long long i = i0; i < m; i += stride) u[i] += (float)alpha * (float)ap[i];
  }
  double t1 = dot.value(), t2 = mx;
  block_sum_max(t1, t2, lds2);
  if (threadIdx.x == 0) {
    part_rty[blockIdx.x] = t1;
    part_rn[blockIdx.x] = t2;
  }
}

// head of CG iteration j >= 1: r'y and ||r|| from the update's partials,
// beta = r'y / r'y_prev, the stop test; while the loop goes on, p = beta p - y
template <typename V>
__global__ void k_head_pupdate(PcgState* st, const double* part_rty, const double* part_rn, int nparts, int j,
                               int max_iter, int guard, int* flags, int n, V* p, const V* y) {
  __shared__ double lds2[2 * kBlock / kWave];
  double* lds = lds2;
  if (!st->active) {
    if (flags && blockIdx.x == 0 && threadIdx.x == 0) flags[j] = 2 * st->solves;
    return;
  }
  // the first grid-stride element is loaded before the partial sums
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  V p0 = 0, y0 = 0;
  if (i0 < n) {
    p0 = p[i0];
    y0 = y[i0];
  }
  double rty, rn;
  parts_sum_max(part_rty, part_rn, nparts, lds2, &rty, &rn);
  const double beta64 = rty / st->rty[(j - 1) & 1];
  const int go = loop_go(st, rty, rn, j, max_iter, guard);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->beta = beta64;
    st->iter = j;
    st->total_iters++;
    st->rty[j & 1] = rty;
    st->rnorm = rn;
    st->active = go;
    st->done = !go;
    if (flags) flags[j] = 2 * st->solves + go;
  }
  if (!go) return;
  const V beta = (V)beta64;
  if (i0 < n) p[i0] = beta * p0 - y0;
  for (long long i = i0 + (long long)gridDim.x * blockDim.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = beta * p[i] - y[i];
}

// diag preconditioner: d = sigma + diag(P) + sum_j rho_j A_ji^2
template <int G>
__global__ __launch_bounds__(kBlock) void k_precond(int n, double sigma, const double* pval, const int* diag_pos,
                                                    const int* trp, const int* tci, const double* tval,
                                                    const double* rho, double rho_sc, double* d) {
  const int lane = threadIdx.x & (G - 1);
  const int rpb = kBlock / G;
  for (int row = blockIdx.x * rpb + threadIdx.x / G; row < n; row += gridDim.x * rpb) {
    double acc = 0.0;
    if (trp) {
      const int e = trp[row + 1];
      for (int k = trp[row] + lane; k < e; k += G) {
        double a = tval[k];
        acc += a * a * rho_at(rho, rho_sc, tci[k]);
      }
    }
    acc = group_sum<G>(acc);
    if (lane == 0) {
      double pd = diag_pos[row] >= 0 ? pval[diag_pos[row]] : 0.0;
      d[row] = sigma + pd + acc;
    }
  }
}

template <typename V>
__global__ void k_recip_to(int n, const double* d, V* dinv, int none) {
  GRID_STRIDE(i, n) dinv[i] = none ? (V)1 : (V)(1.0 / d[i]);
}

__global__ void k_widen(int n, const float* a, double* b) { GRID_STRIDE(i, n) b[i] = (double)a[i]; }

template <typename V>
__global__ void k_cast(long long n, const double* a, V* b) {
  GRID_STRIDE(i, n) b[i] = (V)a[i];
}

__global__ void k_state_init(PcgState* st, double tol_fraction) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    memset(st, 0, sizeof(PcgState));
    st->red_factor = tol_fraction;
    st->eps_prev = 1.0;
    st->scale = 1.0;
  }
}

__global__ void k_set_active(PcgState* st, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->active = v;
    st->done = !v;
  }
}

// ---- the solver object ----------------------------------------------------------
struct HipPcg {
  LinSysSolver base;  // must stay first: LinSysSolver* <-> HipPcg*
  int n, m, bits, polishing, precond_type;
  int max_iter, reduction_threshold;
  double tol_fraction, sigma, rho_sc;
  const OSQPMatrix *P, *A;
  const OSQPVectorf* rho_vec;
  OSQPFloat *scaled_prim_res, *scaled_dual_res;
  // FP64 vectors: x (warm start), rhs, residual, w, cached A x
  double *x64, *rhs64, *r64, *w64, *ax64;
  int ax_valid;       // ax64 == A x64 for the current x64 and A values
  long long ax_Aver;  // A version ax64 was computed with
  // CG vectors in the CG precision (double or float)
  void *d, *p, *Kp, *r, *y, *w, *dinv;
  double* dtmp;       // FP64 n-vector (preconditioner diagonal)
  PcgState* st;
  double* parts;      // 5 * kMaxReduceBlocks partial slots
  PcgState* h_st;     // pinned mirror for counters
  int* h_flags;       // host-coherent per-iteration loop flags
  int* d_flags;       // device address of h_flags
  int flags_len;
  int host_solves;    // solves queued (matches st->solves once they ran)
  int gA, gT, gK;     // lanes per row: A rows, A' rows, K rows
  int lookahead;      // CG iterations queued ahead of the last observed flag
  int last_cg;        // CG iterations of the previous solve (where its loop stopped)
  int w_given;        // w64 = rho o (A x64 - b2) of the next solve already written (host core's fused update)
  // combined K-operator CSR [P A'] + its row blocks: FP64 values for the
  // defect-correction residual, FP32 values for the mixed CG
  int *krp, *kci, *k_rows, k_nb;
  double* kval64;
  float* kval32;
  long long kver_P, kver_A;
  // column-panelled copies of A (m x n) and K (n x (n+m)) for the hot SpMVs:
  // FP32 ones for the mixed CG, FP64 ones for the residual and z = A x (and
  // for everything in FP64 mode).  An operator whose panels do not pay
  // (panel_worth) has none (ntask == 0 / null) and uses the row-block CSR
  // kernels.  The FP64 panels of A are the matrix's own (matrix_panels64,
  // shared with the residual products of the termination checks).
  PanelMatrix pA32, pK32, pK64;
  const PanelMatrix* pA64;
  void *partA, *partK;  // per-panel partial sums (sized for the largest use)
  // FP32 mode: FP32 copies of the FP64 panels' values.  The defect
  // correction's products (r = K x64 - rhs, z = A x64) read them with x, the
  // products and the sums in FP64: 6 instead of 10 bytes per nonzero.  The
  // correction then solves K~ x = rhs for K~ = P~ + sigma I + A~' R A~ built
  // from the FP32-rounded P and A (relative change <= 2^-24 per entry, below
  // the CG's 1e-7 tolerance floor), and z~ = A~ x, as the CG does; the
  // termination checks still use the FP64 matrices.
  float *dcA32, *dcK32;
  long long dcA_ver, dcK_ver;
  // hybrid operators (panels for the dense part, remainder CSR) where plain
  // panels do not pay: FP32 ones for the CG, FP64 ones for the FP64 CG or the
  // defect correction (with FP32 copies of their values in FP32 mode)
  HybridOp hA32, hK32, hA64, hK64;
  // FP32 mode with panelled or hybrid A: z = A x64 from the CG's own products
  // (u = A d accumulated as d is, in FP32; k_dc_apply_incr) on all but every kZRefresh-th
  // solve, which recomputes it exactly from x64.  ap: the last A p (FP32).
  float* ap;
  float* u;
  int z_incr, z_count, z_use;
};

// solves between two exact recomputations of the cached A x64
constexpr int kZRefresh = 8;

template <typename V>
const HybridOp* hyb_A(const HipPcg* s) {
  const HybridOp* H = sizeof(V) == 4 ? &s->hA32 : &s->hA64;
  return H->pan.ntask > 0 ? H : nullptr;
}
template <typename V>
const HybridOp* hyb_K(const HipPcg* s) {
  const HybridOp* H = sizeof(V) == 4 ? &s->hK32 : &s->hK64;
  return H->pan.ntask > 0 ? H : nullptr;
}

// the panel copy of A / K for value type V, or null when CSR kernels are used
template <typename V>
const PanelMatrix* panels_A(const HipPcg* s) {
  const PanelMatrix* P = sizeof(V) == 4 ? &s->pA32 : s->pA64;
  return P && P->ntask > 0 ? P : nullptr;
}
template <typename V>
const PanelMatrix* panels_K(const HipPcg* s) {
  const PanelMatrix* P = sizeof(V) == 4 ? &s->pK32 : &s->pK64;
  return P->ntask > 0 ? P : nullptr;
}

void refresh_fp32(const OSQPMatrix* Mc) {
  OSQPMatrix* M = (OSQPMatrix*)Mc;
  if (!M->fp32_dirty) return;
  if (!M->val32) M->val32 = (float*)dev_alloc(sizeof(float) * (size_t)(M->nnz > 0 ? M->nnz : 1));
  if (M->nnz > 0)
    hipLaunchKernelGGL(k_cast<float>, dim3(grid_for(M->nnz, kBlock, 16384)), dim3(kBlock), 0, S(), (long long)M->nnz,
                       (const double*)M->val, M->val32);
  M->fp32_dirty = 0;
}

void fill_kvals(HipPcg* s, int cols_too) {
  const int n = s->n;
  const int* trp = s->m ? s->A->trp : nullptr;
  int grid = grid_for((long long)n * kWave, kBlock, 16384);
  hipLaunchKernelGGL(k_kfill<double>, dim3(grid), dim3(kBlock), 0, S(), n, (const int*)s->P->rp,
                     (const int*)s->P->ci, (const double*)s->P->val, trp, (const int*)s->A->tci,
                     (const double*)s->A->tval, (const int*)s->krp, s->kci, s->kval64, cols_too);
  if (s->kval32)
    hipLaunchKernelGGL(k_kfill<float>, dim3(grid), dim3(kBlock), 0, S(), n, (const int*)s->P->rp,
                       (const int*)s->P->ci, (const double*)s->P->val, trp, (const int*)s->A->tci,
                       (const double*)s->A->tval, (const int*)s->krp, s->kci, s->kval32, 0);
  s->kver_P = s->P->version;
  s->kver_A = s->A->version;
}

void refresh_kvals(HipPcg* s) {
  if (s->kver_P == s->P->version && s->kver_A == s->A->version) return;
  fill_kvals(s, 0);
}

void refresh_panels(HipPcg* s) {
  PanelMatrix* Ks[2] = {&s->pK32, &s->pK64};
  if (s->pA32.ntask > 0 && s->pA32.ver != s->A->version) {
    panel_refresh_values(&s->pA32, s->A->val);
    s->pA32.ver = s->A->version;
  }
  if (s->pA64) s->pA64 = matrix_panels64(s->A, 0);  // refreshes the values
  const long long kv = s->P->version * 1000003LL + s->A->version;
  for (PanelMatrix* P : Ks)
    if (P->ntask > 0 && P->ver != kv) {
      panel_refresh_values(P, s->kval64);
      P->ver = kv;
    }
  if (s->dcA32 && s->pA64 && s->dcA_ver != s->A->version) {
    panel_fill_f32vals(*s->pA64, s->A->val, s->dcA32);
    s->dcA_ver = s->A->version;
  }
  if (s->dcK32 && s->dcK_ver != kv) {
    panel_fill_f32vals(s->pK64, s->kval64, s->dcK32);
    s->dcK_ver = kv;
  }
  HybridOp* HA[2] = {&s->hA32, &s->hA64};
  for (HybridOp* H : HA)
    if (H->pan.ntask > 0 && H->ver != s->A->version) {
      hybrid_refresh_values(H, s->A->val);
      H->ver = s->A->version;
    }
  HybridOp* HK[2] = {&s->hK32, &s->hK64};
  for (HybridOp* H : HK)
    if (H->pan.ntask > 0 && H->ver != kv) {
      hybrid_refresh_values(H, s->kval64);
      H->ver = kv;
    }
}

// The FP64-vector products over FP64-built panels: FP32 values in FP32 mode
// (the defect correction), FP64 values otherwise.
void dc_panel_spmv(HipPcg* s, const PanelMatrix& P, const int* active, const double* x0, int n0, const double* x1,
                   double* partial) {
  const float* v32 = &P == s->pA64 ? s->dcA32 : (&P == &s->pK64 ? s->dcK32 : nullptr);
  if (v32) panel_spmv_f32vals(P, v32, active, x0, n0, x1, partial);
  else panel_spmv(P, active, x0, n0, x1, partial);
}
// algorithmic bytes of a hybrid operator application without its epilogue:
// the panel SpMV (values of pv bytes, vectors and partials of xb bytes), the
// partials read back, and the remainder CSR (values, 32-bit columns, row
// pointers, one gathered element per entry)
double hyb_bytes(const HybridOp& H, double pv, double xb) {
  const PanelMatrix& P = H.pan;
  const double rv = H.rv32 && (pv == 4.0 || !H.rv64) ? 4.0 : 8.0;
  return panel_spmv_bytes(P, (int)xb) - (xb - pv) * (double)P.nnz + (double)P.np * H.Rp * xb +
         (double)H.rnnz * (rv + 4.0 + xb) + (H.R + 1.0) * 4.0;
}
// value bytes of a hybrid's panel SpMV for vectors of xb bytes
double hyb_pv(const HybridOp& H, double xb) { return xb == 8.0 && H.dc32 ? 4.0 : xb; }

// algorithmic bytes of one dc_panel_spmv launch: values of 4 or 8 bytes, the
// rest as panel_spmv_bytes with FP64 x and partials
double dc_panel_bytes(const HipPcg* s, const PanelMatrix& P) {
  const bool f32 = (&P == s->pA64 && s->dcA32) || (&P == &s->pK64 && s->dcK32);
  return panel_spmv_bytes(P, 8) - (f32 ? 4.0 * (double)P.nnz : 0.0);
}

int build_panels(HipPcg* s) {
  const int n = s->n, m = s->m;
  const long long nnzA = m ? (long long)s->A->nnz_in : 0;
  const long long knnz = (long long)s->P->nnz + nnzA;
  const int prec[2] = {32, 64};
  bool needK = false, needA = false;
  for (int bits : prec) {
    if (bits == 32 && s->bits != 32) continue;
    needK = needK || panel_worth(knnz, n, n + m, bits);
    needA = needA || (m && panel_worth(nnzA, m, n, bits));
  }
  size_t bytesA = 0, bytesK = 0;
  if (needK) {
    std::vector<int> hrp((size_t)n + 1), hci((size_t)(knnz > 0 ? knnz : 1));
    download(hrp.data(), s->krp, sizeof(int) * hrp.size());
    if (knnz) download(hci.data(), s->kci, sizeof(int) * (size_t)knnz);
    for (int bits : prec) {
      if ((bits == 32 && s->bits != 32) || !panel_worth(knnz, n, n + m, bits)) continue;
      PanelMatrix* P = bits == 32 ? &s->pK32 : &s->pK64;
      if (panel_build(P, n, n + m, bits, hrp.data(), hci.data())) return 1;
      const size_t b = (size_t)P->np * (size_t)n * (bits / 8);
      bytesK = b > bytesK ? b : bytesK;
    }
  }
  if (needA && s->bits == 32 && panel_worth(nnzA, m, n, 32)) {
    std::vector<int> arp((size_t)m + 1), aci((size_t)(nnzA > 0 ? nnzA : 1));
    download(arp.data(), s->A->rp, sizeof(int) * arp.size());
    if (nnzA) download(aci.data(), s->A->ci, sizeof(int) * (size_t)nnzA);
    if (panel_build(&s->pA32, m, n, 32, arp.data(), aci.data())) return 1;
    bytesA = (size_t)s->pA32.np * (size_t)m * sizeof(float);
  }
  if (needA && panel_worth(nnzA, m, n, 64)) {
    s->pA64 = matrix_panels64(s->A, 0);
    if (!s->pA64) return 1;
    const size_t b = (size_t)s->pA64->np * (size_t)m * sizeof(double);
    bytesA = b > bytesA ? b : bytesA;
  }
  // hybrids where plain panels do not pay
  for (int bits : prec) {
    if (bits == 32 && s->bits != 32) continue;
    const bool haveK = bits == 32 ? s->pK32.ntask > 0 : s->pK64.ntask > 0;
    const bool haveA = bits == 32 ? s->pA32.ntask > 0 : (s->pA64 && s->pA64->ntask > 0);
    if (!haveK && n > 0) {
      std::vector<int> hrp((size_t)n + 1), hci((size_t)(knnz > 0 ? knnz : 1));
      download(hrp.data(), s->krp, sizeof(int) * hrp.size());
      if (knnz) download(hci.data(), s->kci, sizeof(int) * (size_t)knnz);
      HybridOp* H = bits == 32 ? &s->hK32 : &s->hK64;
      const int rc = hybrid_build(H, n, n + m, bits, hrp.data(), hci.data());
      if (rc < 0 || (rc == 0 && bits == 64 && s->bits == 32 && hybrid_add_f32(H))) return 1;
      if (rc == 0) {
        const size_t b = (size_t)H->pan.np * (size_t)H->Rp * (bits / 8);
        bytesK = b > bytesK ? b : bytesK;
      }
    }
    if (!haveA && m > 0 && nnzA > 0) {
      std::vector<int> arp((size_t)m + 1), aci((size_t)nnzA);
      download(arp.data(), s->A->rp, sizeof(int) * arp.size());
      download(aci.data(), s->A->ci, sizeof(int) * (size_t)nnzA);
      HybridOp* H = bits == 32 ? &s->hA32 : &s->hA64;
      const int rc = hybrid_build(H, m, n, bits, arp.data(), aci.data());
      if (rc < 0 || (rc == 0 && bits == 64 && s->bits == 32 && hybrid_add_f32(H))) return 1;
      if (rc == 0) {
        const size_t b = (size_t)H->pan.np * (size_t)H->Rp * (bits / 8);
        bytesA = b > bytesA ? b : bytesA;
      }
    }
  }
  if (bytesA && !(s->partA = dev_alloc(bytesA))) return 1;
  if (bytesK && !(s->partK = dev_alloc(bytesK))) return 1;
  s->dcA_ver = s->dcK_ver = -1;
  if (s->bits == 32 && m > 0 && (panels_A<float>(s) || hyb_A<float>(s))) {
    s->ap = (float*)dev_alloc(sizeof(float) * (size_t)m);
    s->u = (float*)dev_alloc(sizeof(float) * (size_t)m);
    if (!s->ap || !s->u) return 1;
    s->z_incr = 1;
  }
  if (s->bits == 32) {
    if (s->pA64 && s->pA64->ntask > 0 && !(s->dcA32 = panel_alloc_f32vals(*s->pA64))) return 1;
    if (s->pK64.ntask > 0 && !(s->dcK32 = panel_alloc_f32vals(s->pK64))) return 1;
  }
  refresh_panels(s);
  return 0;
}

int build_kcsr(HipPcg* s) {
  const int n = s->n;
  const int* trp = s->m ? s->A->trp : nullptr;
  const long long knnz = (long long)s->P->nnz + (s->m ? (long long)s->A->nnz_in : 0);
  s->krp = (int*)dev_alloc(sizeof(int) * (size_t)(n + 1));
  s->kci = (int*)dev_alloc(sizeof(int) * (size_t)(knnz > 0 ? knnz : 1));
  s->kval64 = (double*)dev_alloc(sizeof(double) * (size_t)(knnz > 0 ? knnz : 1));
  s->kval32 = s->bits == 32 ? (float*)dev_alloc(sizeof(float) * (size_t)(knnz > 0 ? knnz : 1)) : nullptr;
  if (!s->krp || !s->kci || !s->kval64 || (s->bits == 32 && !s->kval32)) return 1;
  hipLaunchKernelGGL(k_krp, dim3(grid_for(n + 1, kBlock, 4096)), dim3(kBlock), 0, S(), n, (const int*)s->P->rp,
                     trp, s->krp);
  fill_kvals(s, 1);
  // row blocks of K (host partition of the downloaded row pointer)
  int* hk = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  int* hb = (int*)malloc(sizeof(int) * (size_t)(n + 2));
  if (!hk || !hb) {
    free(hk);
    free(hb);
    return 1;
  }
  download(hk, s->krp, sizeof(int) * (size_t)(n + 1));
  s->k_nb = build_row_blocks(hk, n, hb);
  s->k_rows = (int*)dev_alloc(sizeof(int) * (size_t)(s->k_nb + 1));
  upload(s->k_rows, hb, sizeof(int) * (size_t)(s->k_nb + 1));
  free(hk);
  free(hb);
  return s->k_rows ? 0 : 1;
}

void update_precond(HipPcg* s) {
  if (s->bits == 32) refresh_fp32(s->A);
  refresh_kvals(s);
  refresh_panels(s);
  const double* rho = s->rho_vec ? s->rho_vec->d : nullptr;
  const int G = s->gT;
  int grid = rows_grid(s->n, G, 16384);
  const int* trp = s->m ? s->A->trp : nullptr;
  const double* pval = s->P->val;
  const int* dpos = s->P->diag_pos;
  const int* tci = s->A->tci;
  const double* tval = s->A->tval;
  switch (G) {
    case 4: hipLaunchKernelGGL(k_precond<4>, dim3(grid), dim3(kBlock), 0, S(), s->n, s->sigma, pval, dpos, trp, tci, tval, rho, s->rho_sc, s->dtmp); break;
    case 8: hipLaunchKernelGGL(k_precond<8>, dim3(grid), dim3(kBlock), 0, S(), s->n, s->sigma, pval, dpos, trp, tci, tval, rho, s->rho_sc, s->dtmp); break;
    case 16: hipLaunchKernelGGL(k_precond<16>, dim3(grid), dim3(kBlock), 0, S(), s->n, s->sigma, pval, dpos, trp, tci, tval, rho, s->rho_sc, s->dtmp); break;
    case 32: hipLaunchKernelGGL(k_precond<32>, dim3(grid), dim3(kBlock), 0, S(), s->n, s->sigma, pval, dpos, trp, tci, tval, rho, s->rho_sc, s->dtmp); break;
    default: hipLaunchKernelGGL(k_precond<64>, dim3(grid), dim3(kBlock), 0, S(), s->n, s->sigma, pval, dpos, trp, tci, tval, rho, s->rho_sc, s->dtmp); break;
  }
  int none = s->precond_type == OSQP_NO_PRECONDITIONER ? 1 : 0;
  if (s->bits == 32)
    hipLaunchKernelGGL(k_recip_to<float>, dim3(grid_for(s->n, kBlock, 4096)), dim3(kBlock), 0, S(), s->n,
                       (const double*)s->dtmp, (float*)s->dinv, none);
  else
    hipLaunchKernelGGL(k_recip_to<double>, dim3(grid_for(s->n, kBlock, 4096)), dim3(kBlock), 0, S(), s->n,
                       (const double*)s->dtmp, (double*)s->dinv, none);
  LAUNCHCHK();
}

// Host-side stop rule: before queueing iteration `it`, look at the flag of
// iteration it - lookahead.  The loop-head kernel of that iteration writes
// 2 * solve + active; while it still holds an older solve's value the GPU has
// not got there yet and the host waits (yielding), bounded by a deadline after
// which it simply keeps queueing (the kernels exit on the device flag anyway).
// Iteration 0 is always queued.  Short iterations (small operators) run two
// ahead; a skipped iteration costs six launches, less than waiting for the
// host.  Long ones use lookahead 0 with one launch queued ahead (solve_impl).
// The answer is definitive: past the deadline the host waits for the stream
// and reads the flag then (a flag never written counts as a stop).
bool host_should_stop(HipPcg* s, int it) {
  if (!s->h_flags || it < (s->lookahead > 0 ? s->lookahead : 1)) return false;
  volatile int* f = (volatile int*)s->h_flags;
  const int want = 2 * s->host_solves;
  struct timespec t0, t;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  // busy polling for the first millisecond (the answer usually comes within
  // microseconds), yielding the CPU after that
  for (long k = 1;; ++k) {
    int v = f[it - s->lookahead];
    if (v >= want) return (v & 1) == 0;
    if ((k & 255) == 0) {
      clock_gettime(CLOCK_MONOTONIC, &t);
      const double el = (t.tv_sec - t0.tv_sec) + 1e-9 * (t.tv_nsec - t0.tv_nsec);
      if (el > 5.0) break;
      if (el > 1e-3) sched_yield();
    }
  }
  HIPCHK(hipStreamSynchronize(S()));
  const int v = f[it - s->lookahead];
  return v < want || (v & 1) == 0;
}

// ---- operators: panel SpMV + combine, or the row-block CSR kernel ----------
template <typename V>
const V* csr_vals_A(HipPcg* s) {
  return (const V*)(sizeof(V) == 4 ? (const void*)s->A->val32 : (const void*)s->A->val);
}
template <typename V>
const V* csr_vals_K(HipPcg* s) {
  return (const V*)(sizeof(V) == 4 ? (const void*)s->kval32 : (const void*)s->kval64);
}
const int* active_flag(HipPcg* s) { return (const int*)((const char*)s->st + offsetof(PcgState, active)); }

// ---- hybrid dispatch: value types and lane-group sizes ------------------------
// The panel SpMV of a hybrid: FP32 values for the FP32 CG; for FP64 vectors
// the FP32 copies where the hybrid has them (FP32 mode: defect correction),
// else the FP64 values.
template <typename V>
void hyb_panel(const HybridOp& H, const int* active, const V* x0, int n0, const V* x1, V* partial) {
  if constexpr (sizeof(V) == 4) panel_spmv(H.pan, active, x0, n0, x1, partial);
  else if (H.dc32) panel_spmv_f32vals(H.pan, H.dc32, active, x0, n0, x1, partial);
  else panel_spmv(H.pan, active, x0, n0, x1, partial);
}
// f(Rem<TR>) with the remainder values that go with hyb_panel's
template <typename V, typename F>
void with_rem(const HybridOp& H, F f) {
  if constexpr (sizeof(V) == 4) f(Rem<float>{H.Rp, H.rrp, H.rci, H.rv32});
  else if (H.dc32) f(Rem<float>{H.Rp, H.rrp, H.rci, H.rv32});
  else f(Rem<double>{H.Rp, H.rrp, H.rci, H.rv64});
}
template <typename F>
void with_G(int G, F f) {
  if (G == 8) f(std::integral_constant<int, 8>());
  else if (G == 4) f(std::integral_constant<int, 4>());
  else if (G == 2) f(std::integral_constant<int, 2>());
  else f(std::integral_constant<int, 1>());
}
template <typename R>
using rem_t = std::remove_const_t<std::remove_pointer_t<decltype(R::val)>>;

// w = rho o (A v)  (skipped while the CG state is inactive), in two launches
// where A is panelled (or hybrid): the panel SpMV (av_spmv) and the combine
// (av_finish); the row-block CSR kernel does both in av_spmv.
template <typename V>
void av_spmv(HipPcg* s, const V* v, V* w) {
  const double* rho = s->rho_vec ? s->rho_vec->d : nullptr;
  if (const PanelMatrix* P = panels_A<V>(s)) {
    if constexpr (sizeof(V) == 8) dc_panel_spmv(s, *P, active_flag(s), v, s->n, nullptr, (double*)s->partA);
    else panel_spmv(*P, active_flag(s), v, s->n, (const V*)nullptr, (V*)s->partA);
  } else if (const HybridOp* H = hyb_A<V>(s)) {
    hyb_panel<V>(*H, active_flag(s), v, s->n, (const V*)nullptr, (V*)s->partA);
  } else {
    const int g = s->A->rb_n < kRowGridCap ? s->A->rb_n : kRowGridCap;
    hipLaunchKernelGGL((k_rb_Av_rho<V, V>), dim3(g), dim3(kBlock), 0, S(), (const PcgState*)s->st,
                       (const int*)s->A->rb_rows, s->A->rb_n, (const int*)s->A->rp, (const int*)s->A->ci,
                       csr_vals_A<V>(s), v, rho, s->rho_sc, w);
  }
}
template <typename V>
void av_finish(HipPcg* s, const V* v, V* w) {
  const int m = s->m;
  const double* rho = s->rho_vec ? s->rho_vec->d : nullptr;
  if (const HybridOp* H = panels_A<V>(s) ? nullptr : hyb_A<V>(s)) {
    const int G = comb_group(H->pan.np, m);
    const dim3 g(grid_for((long long)m * G, kBlock, 8192));
    V* ap = sizeof(V) == 4 && s->z_use ? (V*)s->ap : nullptr;
    with_rem<V>(*H, [&](auto rm) {
      with_G(G, [&](auto Gc) {
        hipLaunchKernelGGL((k_hyb_av<V, decltype(Gc)::value, rem_t<decltype(rm)>>), g, dim3(kBlock), 0, S(),
                           (const PcgState*)s->st, m, H->pan.np, (const V*)s->partA, rm, v, rho, s->rho_sc, w, ap);
      });
    });
    return;
  }
  const PanelMatrix* P = panels_A<V>(s);
  if (!P) return;
  V* ap = sizeof(V) == 4 && s->z_use ? (V*)s->ap : nullptr;
  const int G = comb_group(P->np, m);
  const dim3 g(grid_for((long long)m * G, kBlock, 8192));
  if (G == 8)
    hipLaunchKernelGGL((k_comb_av<V, 8>), g, dim3(kBlock), 0, S(), (const PcgState*)s->st, m, P->np,
                       (const V*)s->partA, rho, s->rho_sc, w, ap);
  else if (G == 4)
    hipLaunchKernelGGL((k_comb_av<V, 4>), g, dim3(kBlock), 0, S(), (const PcgState*)s->st, m, P->np,
                       (const V*)s->partA, rho, s->rho_sc, w, ap);
  else if (G == 2)
    hipLaunchKernelGGL((k_comb_av<V, 2>), g, dim3(kBlock), 0, S(), (const PcgState*)s->st, m, P->np,
                       (const V*)s->partA, rho, s->rho_sc, w, ap);
  else
    hipLaunchKernelGGL((k_comb_av<V, 1>), g, dim3(kBlock), 0, S(), (const PcgState*)s->st, m, P->np,
                       (const V*)s->partA, rho, s->rho_sc, w, ap);
}
template <typename V>
void op_av_rho(HipPcg* s, const V* v, V* w) {
  av_spmv<V>(s, v, w);
  av_finish<V>(s, v, w);
}

// Kv = sigma v + [P A'] [v; w] (- sub) and block partials (v'Kv, or
// max|Kv - sub| when sub is given); returns the number of partials
template <typename V>
int op_kv(HipPcg* s, const V* v, const V* w, const double* sub, V* Kv, double* parts) {
  const int n = s->n;
  if (const PanelMatrix* P = panels_K<V>(s)) {
    if constexpr (sizeof(V) == 8) dc_panel_spmv(s, *P, sub ? nullptr : active_flag(s), v, n, w, (double*)s->partK);
    else panel_spmv(*P, sub ? nullptr : active_flag(s), v, n, w, (V*)s->partK);
    constexpr int RB = 16 / (int)sizeof(V);
    if (n % RB == 0) {
      const int nq = n / RB;
      const int G = comb_group(P->np, nq);
      const int g = grid_for((long long)nq * G, kBlock, kMaxReduceBlocks);
      if (G == 8)
        hipLaunchKernelGGL((k_comb_kv_rows<V, 8>), dim3(g), dim3(kBlock), 0, S(), (const PcgState*)s->st, n, P->np,
                           (const V*)s->partK, s->sigma, v, sub, Kv, parts);
      else if (G == 4)
        hipLaunchKernelGGL((k_comb_kv_rows<V, 4>), dim3(g), dim3(kBlock), 0, S(), (const PcgState*)s->st, n, P->np,
                           (const V*)s->partK, s->sigma, v, sub, Kv, parts);
      else if (G == 2)
        hipLaunchKernelGGL((k_comb_kv_rows<V, 2>), dim3(g), dim3(kBlock), 0, S(), (const PcgState*)s->st, n, P->np,
                           (const V*)s->partK, s->sigma, v, sub, Kv, parts);
      else
        hipLaunchKernelGGL((k_comb_kv_rows<V, 1>), dim3(g), dim3(kBlock), 0, S(), (const PcgState*)s->st, n, P->np,
                           (const V*)s->partK, s->sigma, v, sub, Kv, parts);
      return g;
    }
    const int G = comb_group(P->np, n);
    const int g = grid_for((long long)n * G, kBlock, kMaxReduceBlocks);
    if (G == 8)
      hipLaunchKernelGGL((k_comb_kv<V, 8>), dim3(g), dim3(kBlock), 0, S(), (const PcgState*)s->st, n, P->np,
                         (const V*)s->partK, s->sigma, v, sub, Kv, parts);
    else if (G == 4)
      hipLaunchKernelGGL((k_comb_kv<V, 4>), dim3(g), dim3(kBlock), 0, S(), (const PcgState*)s->st, n, P->np,
                         (const V*)s->partK, s->sigma, v, sub, Kv, parts);
    else if (G == 2)
      hipLaunchKernelGGL((k_comb_kv<V, 2>), dim3(g), dim3(kBlock), 0, S(), (const PcgState*)s->st, n, P->np,
                         (const V*)s->partK, s->sigma, v, sub, Kv, parts);
    else
      hipLaunchKernelGGL((k_comb_kv<V, 1>), dim3(g), dim3(kBlock), 0, S(), (const PcgState*)s->st, n, P->np,
                         (const V*)s->partK, s->sigma, v, sub, Kv, parts);
    return g;
  }
  if (const HybridOp* H = hyb_K<V>(s)) {
    hyb_panel<V>(*H, sub ? nullptr : active_flag(s), v, n, w, (V*)s->partK);
    const int G = comb_group(H->pan.np, n);
    const int g = grid_for((long long)n * G, kBlock, kMaxReduceBlocks);
    with_rem<V>(*H, [&](auto rm) {
      with_G(G, [&](auto Gc) {
        hipLaunchKernelGGL((k_hyb_kv<V, decltype(Gc)::value, rem_t<decltype(rm)>>), dim3(g), dim3(kBlock), 0, S(),
                           (const PcgState*)s->st, n, H->pan.np, (const V*)s->partK, rm, s->sigma, v, w, sub, Kv,
                           parts);
      });
    });
    return g;
  }
  const int g = s->k_nb < kMaxReduceBlocks ? s->k_nb : kMaxReduceBlocks;
  hipLaunchKernelGGL((k_rb_Kv<V, V>), dim3(g), dim3(kBlock), 0, S(), (const PcgState*)s->st, (const int*)s->k_rows,
                     s->k_nb, (const int*)s->krp, (const int*)s->kci, csr_vals_K<V>(s), n, s->sigma, v, w, sub, Kv,
                     parts);
  return g;
}

// b[n:n+m] = A x64 (FP64 values), the A x cache; polishing: rho_pol (A x - b2).
// With `when` (panelled A only) the launches do nothing while *when == 0.
void op_finish_z(HipPcg* s, double* b, const int* when) {
  const int n = s->n, m = s->m;
  if (const PanelMatrix* P = panels_A<double>(s)) {
    dc_panel_spmv(s, *P, when, (const double*)s->x64, n, nullptr, (double*)s->partA);
    const int G = comb_group(P->np, m);
    const dim3 g(grid_for((long long)m * G, kBlock, 8192));
    if (G == 8)
      hipLaunchKernelGGL(k_comb_z<8>, g, dim3(kBlock), 0, S(), when, m, n, P->np, (const double*)s->partA, b,
                         s->ax64, s->polishing, s->rho_sc);
    else if (G == 4)
      hipLaunchKernelGGL(k_comb_z<4>, g, dim3(kBlock), 0, S(), when, m, n, P->np, (const double*)s->partA, b,
                         s->ax64, s->polishing, s->rho_sc);
    else if (G == 2)
      hipLaunchKernelGGL(k_comb_z<2>, g, dim3(kBlock), 0, S(), when, m, n, P->np, (const double*)s->partA, b,
                         s->ax64, s->polishing, s->rho_sc);
    else
      hipLaunchKernelGGL(k_comb_z<1>, g, dim3(kBlock), 0, S(), when, m, n, P->np, (const double*)s->partA, b,
                         s->ax64, s->polishing, s->rho_sc);
  } else if (const HybridOp* H = hyb_A<double>(s)) {
    hyb_panel<double>(*H, when, (const double*)s->x64, n, (const double*)nullptr, (double*)s->partA);
    const int G = comb_group(H->pan.np, m);
    const dim3 g(grid_for((long long)m * G, kBlock, 8192));
    with_rem<double>(*H, [&](auto rm) {
      with_G(G, [&](auto Gc) {
        hipLaunchKernelGGL((k_hyb_z<decltype(Gc)::value, rem_t<decltype(rm)>>), g, dim3(kBlock), 0, S(), when, m, n,
                           H->pan.np, (const double*)s->partA, rm, (const double*)s->x64, b, s->ax64, s->polishing,
                           s->rho_sc);
      });
    });
  } else {
    if (when) {
      check(hipErrorInvalidValue, "op_finish_z: a guarded tail needs panels", __FILE__, __LINE__);
      return;
    }
    const int g = s->A->rb_n < kRowGridCap ? s->A->rb_n : kRowGridCap;
    hipLaunchKernelGGL(k_rb_finish_z<double>, dim3(g), dim3(kBlock), 0, S(), (const int*)s->A->rb_rows, s->A->rb_n,
                       (const int*)s->A->rp, (const int*)s->A->ci, (const double*)s->A->val, (const double*)s->x64,
                       b, s->ax64, n, s->polishing, s->rho_sc);
  }
}

template <typename V>
void solve_impl(HipPcg* s, OSQPVectorf* b, int admm_iter) {
  const int n = s->n, m = s->m;
  const double* rho = s->rho_vec ? s->rho_vec->d : nullptr;
  double* part0 = s->parts;
  double* part1 = s->parts + kMaxReduceBlocks;
  double* part2 = s->parts + 2 * kMaxReduceBlocks;
  double* part3 = s->parts + 3 * kMaxReduceBlocks;  // the init's r'y / max|r| partials
  double* part4 = s->parts + 4 * kMaxReduceBlocks;
  const int gK = s->k_nb < kMaxReduceBlocks ? s->k_nb : kMaxReduceBlocks;
  const int gE = grid_for(n, kBlock, kMaxReduceBlocks);
  const int gM = grid_for(m, kBlock, 4096);
  const int guard = s->bits == 32 ? 1 : 0;
  V* p = (V*)s->p;
  V* Kp = (V*)s->Kp;
  V* r = (V*)s->r;
  V* y = (V*)s->y;
  V* w = (V*)s->w;
  V* d = (V*)s->d;
  const V* dinv = (const V*)s->dinv;

  s->host_solves++;
  // the residual norms steer the tolerance of ADMM solves; a polishing solver
  // has none (its tolerance follows ||rhs||)
  const double spr = s->scaled_prim_res ? *s->scaled_prim_res : 0.0;
  const double sdr = s->scaled_dual_res ? *s->scaled_dual_res : 0.0;
  const bool cached = m > 0 && s->ax_valid && s->ax_Aver == s->A->version && admm_iter > 1 && !s->polishing;
  const bool w_given = cached && s->w_given;
  // incremental z = A x64 on a cached solve, exact every kZRefresh-th one
  s->z_use = s->z_incr && cached && s->z_count < kZRefresh - 1 && sizeof(V) == 4;
  s->z_count = s->z_use ? s->z_count + 1 : 0;
  s->w_given = 0;
  int nres;
  if (!cached) {
    // explicit rhs: needed for ||rhs|| at the first ADMM iteration and in polishing
    hipLaunchKernelGGL(k_rb_rhs<double>, dim3(gK), dim3(kBlock), 0, S(), (const int*)s->k_rows, s->k_nb,
                       (const int*)s->krp, (const int*)s->kci, (const double*)s->kval64, n, (const double*)b->d, rho,
                       s->rho_sc, s->rhs64, part0);
    hipLaunchKernelGGL(k_tolerance, dim3(1), dim3(kBlock), 0, S(), s->st, (const double*)part0, gK, admm_iter,
                       spr, sdr, s->polishing, s->tol_fraction,
                       s->reduction_threshold);
    // r64 = K x64 - rhs64 with the FP64 values (the exact K)
    if (m) op_av_rho<double>(s, (const double*)s->x64, s->w64);
    nres = op_kv<double>(s, (const double*)s->x64, (const double*)s->w64, (const double*)s->rhs64, s->r64, part2);
  } else {
    // tolerance; w64 = rho o (A x64 - b2) from the cached A x64;
    // r64 = sigma x64 + P x64 + A' w64 - b1.  When the host core's fused
    // ADMM update has already written w64 (osqp_amd_linsys_set_w_given), the
    // tolerance update moves into the CG init.
    if (!w_given)
      hipLaunchKernelGGL(k_tol_w, dim3(gM), dim3(kBlock), 0, S(), s->st, admm_iter, spr, sdr, s->tol_fraction,
                         s->reduction_threshold, m, (const double*)s->ax64, (const double*)(b->d + n), rho,
                         s->rho_sc, s->w64);
    nres = op_kv<double>(s, (const double*)s->x64, (const double*)s->w64, (const double*)b->d, s->r64, part2);
  }
  hipLaunchKernelGGL(k_dc_init_cg<V>, dim3(gE), dim3(kBlock), 0, S(), s->st, (const double*)part2, nres, n,
                     (const double*)s->r64, dinv, r, y, p, d, part3, part4, w_given ? 1 : 0, admm_iter, spr, sdr,
                     s->tol_fraction, s->reduction_threshold);
  // one CG iteration: the operator (two SpMVs with their combines), the
  // update, and the next loop head fused with the search-direction update.
  // With lookahead 0 the first launch of iteration it + 1 (A's panel SpMV) is
  // queued before the host waits for iteration it + 1's go, so the GPU has
  // that SpMV to run while the host reacts; a stop skips just that launch.
  //
  // The solve's tail (x64 += scale d, then z = A x64) may be queued before
  // the host knows the loop has stopped: when iteration it + 1 is where the
  // previous solve's loop stopped, the tail is queued in place of the
  // look-ahead SpMV, each launch guarded on the device by the loop's `done`
  // flag.  If the loop stops there, the GPU runs the tail while the host is
  // still reacting; if it goes on, the three launches exit at once, the host
  // queues the SpMV it left out, and the tail is queued unguarded after the
  // loop.  host_should_stop's answers are definitive, so the host knows which
  // of the two happened.
  const bool ahead = s->lookahead == 0 && m > 0;
  const bool spec = ahead && s->h_flags && (panels_A<double>(s) != nullptr || hyb_A<double>(s) != nullptr);
  const int* done = (const int*)((const char*)s->st + offsetof(PcgState, done));
  auto tail = [&](bool guarded) {
    if (m && s->z_use) {
      hipLaunchKernelGGL(k_dc_apply_incr, dim3(grid_for((long long)n + m, kBlock, 8192)), dim3(kBlock), 0, S(),
                         (const PcgState*)s->st, guarded ? 1 : 0, n, m, (const float*)d, (const float*)s->u, s->x64,
                         s->ax64, b->d);
      return;
    }
    hipLaunchKernelGGL(k_dc_apply<V>, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, S(), (const PcgState*)s->st,
                       guarded ? 1 : 0, n, (const V*)d, s->x64, b->d);
    if (m) op_finish_z(s, b->d, guarded ? done : nullptr);
  };
  int spec_at = -1, stop_at = s->max_iter;
  if (ahead) av_spmv<V>(s, (const V*)p, w);
  for (int it = 0; it < s->max_iter; ++it) {
    if (host_should_stop(s, it)) {
      stop_at = it;
      break;
    }
    osqp_amd_sec_push(OSQP_AMD_SEC_LINSYS_MVM);
    if (ahead) {
      // the look-ahead SpMV was left out where the loop was predicted to stop
      if (it == spec_at) av_spmv<V>(s, (const V*)p, w);
      av_finish<V>(s, (const V*)p, w);
    } else if (m) {
      op_av_rho<V>(s, (const V*)p, w);
    }
    const int nk = op_kv<V>(s, (const V*)p, (const V*)w, (const double*)nullptr, Kp, part2);
    osqp_amd_sec_pop(OSQP_AMD_SEC_LINSYS_MVM);
    hipLaunchKernelGGL(k_update<V>, dim3(gE), dim3(kBlock), 0, S(), s->st, (const double*)part2, nk, it, guard, n,
                       d, r, y, (const V*)p, (const V*)Kp, dinv, part0, part1, (const double*)part3,
                       (const double*)part4, gE, s->max_iter, s->d_flags, m,
                       s->z_use ? (const V*)s->ap : nullptr, s->z_use ? s->u : nullptr);
    hipLaunchKernelGGL(k_head_pupdate<V>, dim3(gE), dim3(kBlock), 0, S(), s->st, (const double*)part0,
                       (const double*)part1, gE, it + 1, s->max_iter, guard, s->d_flags, n, p, (const V*)y);
    if (ahead && it + 1 < s->max_iter) {
      if (spec && it + 1 == s->last_cg) {
        // predicted stop: the tail instead of the look-ahead SpMV (a skipped
        // SpMV launch would sit between the loop head and the tail)
        tail(true);
        spec_at = it + 1;
      } else {
        av_spmv<V>(s, (const V*)p, w);
      }
    }
  }
  s->last_cg = stop_at;
  if (spec_at < 0 || spec_at != stop_at) tail(false);
  s->z_use = 0;
  if (m) {
    s->ax_valid = 1;
    s->ax_Aver = s->A->version;
  }
  LAUNCHCHK();
}

// ---- vtable entries ------------------------------------------------------------
const char* pcg_name(LinSysSolver* self) {
  HipPcg* s = (HipPcg*)self;
  return s->bits == 32 ? "HIP PCG (FP32 CG, FP64 defect correction) - diagonal preconditioner"
                       : "HIP PCG (FP64) - diagonal preconditioner";
}

OSQPInt pcg_solve(LinSysSolver* self, OSQPVectorf* b, OSQPInt admm_iter) {
  HipPcg* s = (HipPcg*)self;
  if (s->flags_len < s->max_iter + 1) {
    if (s->h_flags) {
      HIPCHK(hipStreamSynchronize(S()));
      (void)hipHostFree(s->h_flags);
    }
    s->h_flags = nullptr;
    s->d_flags = nullptr;
    s->flags_len = 0;
    if (hipHostMalloc((void**)&s->h_flags, sizeof(int) * (size_t)(s->max_iter + 1),
                      hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess &&
        hipHostGetDevicePointer((void**)&s->d_flags, s->h_flags, 0) == hipSuccess) {
      memset(s->h_flags, 0, sizeof(int) * (size_t)(s->max_iter + 1));
      s->flags_len = s->max_iter + 1;
    } else {
      (void)hipGetLastError();
      if (s->h_flags) (void)hipHostFree(s->h_flags);
      s->h_flags = nullptr;
      s->d_flags = nullptr;
    }
  }
  if (s->kver_P != s->P->version || s->kver_A != s->A->version || (s->bits == 32 && s->A->fp32_dirty))
    update_precond(s);
  osqp_amd_sec_push(OSQP_AMD_SEC_LINSYS_SOLVE);
  if (s->bits == 32) {
    solve_impl<float>(s, b, admm_iter);
  } else {
    solve_impl<double>(s, b, admm_iter);
  }
  osqp_amd_sec_pop(OSQP_AMD_SEC_LINSYS_SOLVE);
  return 0;
}

void pcg_update_settings(LinSysSolver* self, const OSQPSettings* st) {
  HipPcg* s = (HipPcg*)self;
  s->max_iter = st->cg_max_iter;
  s->reduction_threshold = st->cg_tol_reduction;
  s->tol_fraction = st->cg_tol_fraction;
  if (s->precond_type != (int)st->cg_precond) {
    s->precond_type = st->cg_precond;
    update_precond(s);
  }
}

// cuda_pcg_interface.cu:292-296 — the CG iterate restarts from x
void pcg_warm_start(LinSysSolver* self, const OSQPVectorf* x) {
  HipPcg* s = (HipPcg*)self;
  if (s->n > 0)
    HIPCHK(hipMemcpyAsync(s->x64, x->d, sizeof(double) * (size_t)s->n, hipMemcpyDeviceToDevice, S()));
  s->ax_valid = 0;
}

void pcg_free(LinSysSolver* self) {
  HipPcg* s = (HipPcg*)self;
  if (!s) return;
  HIPCHK(hipStreamSynchronize(S()));
  dev_free(s->x64); dev_free(s->rhs64); dev_free(s->r64); dev_free(s->w64); dev_free(s->ax64);
  dev_free(s->d); dev_free(s->p); dev_free(s->Kp); dev_free(s->r); dev_free(s->y);
  dev_free(s->w); dev_free(s->dinv); dev_free(s->dtmp);
  dev_free(s->st); dev_free(s->parts);
  dev_free(s->krp); dev_free(s->kci); dev_free(s->kval64); dev_free(s->kval32); dev_free(s->k_rows);
  panel_free(&s->pA32); panel_free(&s->pK32); panel_free(&s->pK64);
  dev_free(s->dcA32); dev_free(s->dcK32);
  hybrid_free(&s->hA32); hybrid_free(&s->hK32); hybrid_free(&s->hA64); hybrid_free(&s->hK64);
  dev_free(s->ap); dev_free(s->u);
  dev_free(s->partA); dev_free(s->partK);
  if (s->h_st) (void)hipHostFree(s->h_st);
  if (s->h_flags) (void)hipHostFree(s->h_flags);
  free(s);
}

OSQPInt pcg_update_matrices(LinSysSolver* self, const OSQPMatrix*, const OSQPInt*, OSQPInt, const OSQPMatrix*,
                            const OSQPInt*, OSQPInt) {
  HipPcg* s = (HipPcg*)self;
  s->ax_valid = 0;
  s->w_given = 0;
  update_precond(s);
  return 0;
}

OSQPInt pcg_update_rho(LinSysSolver* self, const OSQPVectorf*, OSQPFloat rho_sc) {
  HipPcg* s = (HipPcg*)self;
  s->rho_sc = rho_sc;
  s->w_given = 0;
  update_precond(s);
  return 0;
}

}  // namespace

static int g_pcg_bits = 64;

extern "C" {

OSQPInt osqp_amd_set_pcg_precision(OSQPInt bits) {
  if (bits != 32 && bits != 64) return 1;
  g_pcg_bits = bits;
  return 0;
}

OSQPInt osqp_amd_get_pcg_precision(void) { return g_pcg_bits; }

// include/private/lin_alg.h:37 ; cuda_pcg_interface.cu:99-214
OSQPInt osqp_algebra_init_linsys_solver(LinSysSolver** sp, const OSQPMatrix* P, const OSQPMatrix* A,
                                        const OSQPVectorf* rho_vec, const OSQPSettings* settings,
                                        OSQPFloat* scaled_prim_res, OSQPFloat* scaled_dual_res, OSQPInt polishing) {
  if (settings->linsys_solver != OSQP_INDIRECT_SOLVER) return OSQP_LINSYS_SOLVER_INIT_ERROR;
  HipPcg* s = (HipPcg*)calloc(1, sizeof(HipPcg));
  if (!s) return OSQP_LINSYS_SOLVER_INIT_ERROR;
  *sp = &s->base;
  s->base.type = OSQP_INDIRECT_SOLVER;
  s->base.name = pcg_name;
  s->base.solve = pcg_solve;
  s->base.update_settings = pcg_update_settings;
  s->base.warm_start = pcg_warm_start;
  s->base.adjoint_derivative = nullptr;
  s->base.free = pcg_free;
  s->base.update_matrices = pcg_update_matrices;
  s->base.update_rho_vec = pcg_update_rho;
  s->base.nthreads = 1;
  s->n = P->n;
  s->m = A->m;
  s->bits = g_pcg_bits;
  s->polishing = polishing;
  s->precond_type = settings->cg_precond;
  s->max_iter = settings->cg_max_iter;
  s->reduction_threshold = settings->cg_tol_reduction;
  s->tol_fraction = settings->cg_tol_fraction;
  s->P = P;
  s->A = A;
  s->rho_vec = rho_vec;
  s->scaled_prim_res = scaled_prim_res;
  s->scaled_dual_res = scaled_dual_res;
  s->lookahead = 2;
  if (polishing) {
    s->sigma = settings->delta;
    s->rho_sc = 1.0 / settings->delta;
  } else {
    s->sigma = settings->sigma;
    s->rho_sc = settings->rho;
  }
  const size_t vb = s->bits == 32 ? sizeof(float) : sizeof(double);
  const size_t n = (size_t)(s->n > 0 ? s->n : 1), m = (size_t)(s->m > 0 ? s->m : 1);
  s->x64 = (double*)dev_alloc(sizeof(double) * n);
  s->rhs64 = (double*)dev_alloc(sizeof(double) * n);
  s->r64 = (double*)dev_alloc(sizeof(double) * n);
  s->w64 = (double*)dev_alloc(sizeof(double) * m);
  s->ax64 = (double*)dev_alloc(sizeof(double) * m);
  s->d = dev_alloc(vb * n);
  s->p = dev_alloc(vb * n);
  s->Kp = dev_alloc(vb * n);
  s->r = dev_alloc(vb * n);
  s->y = dev_alloc(vb * n);
  s->w = dev_alloc(vb * m);
  s->dinv = dev_alloc(vb * n);
  s->dtmp = (double*)dev_alloc(sizeof(double) * n);
  s->st = (PcgState*)dev_alloc(sizeof(PcgState));
  s->parts = (double*)dev_alloc(sizeof(double) * 5 * kMaxReduceBlocks);
  if (!s->x64 || !s->rhs64 || !s->r64 || !s->w64 || !s->ax64 || !s->d || !s->p || !s->Kp || !s->r || !s->y ||
      !s->w || !s->dinv || !s->dtmp || !s->st || !s->parts) {
    pcg_free(&s->base);
    *sp = nullptr;
    return OSQP_LINSYS_SOLVER_INIT_ERROR;
  }
  HIPCHK(hipHostMalloc((void**)&s->h_st, sizeof(PcgState), hipHostMallocDefault));
  HIPCHK(hipMemsetAsync(s->x64, 0, sizeof(double) * n, S()));
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, S(), s->st, s->tol_fraction);
  s->gA = A->lanes;
  s->gT = A->tlanes;
  // K rows hold a P row and an A' row: size lanes for their sum
  s->gK = lanes_for_rows((long long)P->nnz + (long long)A->nnz_in, s->n);
  if (build_kcsr(s) || build_panels(s)) {
    pcg_free(&s->base);
    *sp = nullptr;
    return OSQP_LINSYS_SOLVER_INIT_ERROR;
  }
  update_precond(s);
  // no speculative CG iterations where one moves more than 64 MB
  double cg_bytes = 0.0;
  osqp_amd_linsys_bytes(&s->base, &cg_bytes, nullptr);
  s->lookahead = cg_bytes > 64e6 ? 0 : 2;
  return 0;
}

// Times the two CG operator kernels (w = rho o A p and Kp = sigma p + P p + A' w)
// on this solver's matrices: `reps` launches of each, bracketed by HIP events
// on the backend stream.  Reports the mean duration of one launch and the
// algorithmic bytes it moves (values + indices + row pointers, each vector
// element touched once).
OSQPInt osqp_amd_linsys_time_kernels(LinSysSolver* ls, OSQPInt reps, double* ms_av, double* ms_kv,
                                     double* bytes_av, double* bytes_kv) {
  HipPcg* s = (HipPcg*)ls;
  if (!s || reps <= 0) return 1;
  const int n = s->n, m = s->m;
  const double vb = s->bits == 32 ? 4.0 : 8.0;
  hipEvent_t e0, e1, e2;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventCreate(&e2));
  update_precond(s);
  // inputs: p = 1, w = 1 (finite values; the timing does not depend on them)
  if (s->bits == 32) {
    hipLaunchKernelGGL(k_recip_to<float>, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, S(), n,
                       (const double*)s->dtmp, (float*)s->p, 1);
    if (m) hipLaunchKernelGGL(k_recip_to<float>, dim3(grid_for(m, kBlock, 4096)), dim3(kBlock), 0, S(), m,
                              (const double*)s->w64, (float*)s->w, 1);
  } else {
    hipLaunchKernelGGL(k_recip_to<double>, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, S(), n,
                       (const double*)s->dtmp, (double*)s->p, 1);
    if (m) hipLaunchKernelGGL(k_recip_to<double>, dim3(grid_for(m, kBlock, 4096)), dim3(kBlock), 0, S(), m,
                              (const double*)s->w64, (double*)s->w, 1);
  }
  hipLaunchKernelGGL(k_set_active, dim3(1), dim3(64), 0, S(), s->st, 1);
  HIPCHK(hipStreamSynchronize(S()));
  // the SpMV kernels of one CG iteration in the CG precision: the panel
  // kernel alone where A / K are panelled, else the row-block CSR kernel
  const PanelMatrix* PA = s->bits == 32 ? panels_A<float>(s) : panels_A<double>(s);
  const PanelMatrix* PK = s->bits == 32 ? panels_K<float>(s) : panels_K<double>(s);
  HIPCHK(hipEventRecord(e0, S()));
  for (int r = 0; r < reps && m; ++r) {
    if (PA && s->bits == 32)
      panel_spmv(*PA, active_flag(s), (const float*)s->p, n, (const float*)nullptr, (float*)s->partA);
    else if (PA)
      panel_spmv(*PA, active_flag(s), (const double*)s->p, n, (const double*)nullptr, (double*)s->partA);
    else if (s->bits == 32)
      op_av_rho<float>(s, (const float*)s->p, (float*)s->w);
    else
      op_av_rho<double>(s, (const double*)s->p, (double*)s->w);
  }
  HIPCHK(hipEventRecord(e1, S()));
  for (int r = 0; r < reps; ++r) {
    if (PK && s->bits == 32)
      panel_spmv(*PK, active_flag(s), (const float*)s->p, n, (const float*)s->w, (float*)s->partK);
    else if (PK)
      panel_spmv(*PK, active_flag(s), (const double*)s->p, n, (const double*)s->w, (double*)s->partK);
    else if (s->bits == 32)
      op_kv<float>(s, (const float*)s->p, (const float*)s->w, nullptr, (float*)s->Kp, s->parts);
    else
      op_kv<double>(s, (const double*)s->p, (const double*)s->w, nullptr, (double*)s->Kp, s->parts);
  }
  HIPCHK(hipEventRecord(e2, S()));
  HIPCHK(hipEventSynchronize(e2));
  float t01 = 0.f, t12 = 0.f;
  HIPCHK(hipEventElapsedTime(&t01, e0, e1));
  HIPCHK(hipEventElapsedTime(&t12, e1, e2));
  if (ms_av) *ms_av = (double)t01 / reps;
  if (ms_kv) *ms_kv = (double)t12 / reps;
  // bytes per launch.  Panels: values + 16-bit indices, row pointers and one
  // partial per (panel, row), the multiplied vector, task descriptors.  CSR:
  // values + 32-bit indices, row pointers, the vector, the output (+ rho).
  const double nnzA = (double)s->A->nnz_in, nnzK = (double)s->P->nnz + (m ? nnzA : 0.0);
  // Hybrids: the panel SpMV and its combine (hyb_bytes + the epilogue).
  const HybridOp* HA = s->bits == 32 ? hyb_A<float>(s) : hyb_A<double>(s);
  const HybridOp* HK = s->bits == 32 ? hyb_K<float>(s) : hyb_K<double>(s);
  if (bytes_av)
    *bytes_av = PA ? panel_spmv_bytes(*PA, (int)vb)
                   : HA ? hyb_bytes(*HA, vb, vb) + m * (vb + 8.0)
                        : nnzA * (vb + 4.0) + (m + 1.0) * 4.0 + n * vb + m * (vb + 8.0);
  if (bytes_kv)
    *bytes_kv = PK ? panel_spmv_bytes(*PK, (int)vb)
                   : HK ? hyb_bytes(*HK, vb, vb) + 2.0 * n * vb
                        : nnzK * (vb + 4.0) + (n + 1.0) * 4.0 + (n + (double)m) * vb + 2.0 * n * vb;
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  HIPCHK(hipEventDestroy(e2));
  // the CG state of this solver is reset: the next solve starts fresh
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, S(), s->st, s->tol_fraction);
  s->ax_valid = 0;
  s->host_solves = 0;
  HIPCHK(hipStreamSynchronize(S()));
  if (s->h_flags) memset(s->h_flags, 0, sizeof(int) * (size_t)s->flags_len);
  return 0;
}

// Times the FP64 panel SpMVs of the defect correction (z = A x64 and the
// residual's [P A'] [x64; w64]): `reps` launches of each on the backend
// stream.  Returns 1 when the operators are not panelled.
OSQPInt osqp_amd_linsys_time_fp64_kernels(LinSysSolver* ls, OSQPInt reps, double* ms_a, double* ms_k,
                                          double* bytes_a, double* bytes_k) {
  HipPcg* s = (HipPcg*)ls;
  if (!s || reps <= 0) return 1;
  const PanelMatrix* PA = panels_A<double>(s);
  const PanelMatrix* PK = panels_K<double>(s);
  if (!PA || !PK) return 1;
  const int n = s->n;
  hipEvent_t e0, e1, e2;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventCreate(&e2));
  update_precond(s);
  // inputs: x64 (the warm start) and w64 = 1
  hipLaunchKernelGGL(k_recip_to<double>, dim3(grid_for(s->m, kBlock, 4096)), dim3(kBlock), 0, S(), s->m,
                     (const double*)s->dtmp, s->w64, 1);
  HIPCHK(hipStreamSynchronize(S()));
  HIPCHK(hipEventRecord(e0, S()));
  for (int r = 0; r < reps; ++r) dc_panel_spmv(s, *PA, nullptr, (const double*)s->x64, n, nullptr, (double*)s->partA);
  HIPCHK(hipEventRecord(e1, S()));
  for (int r = 0; r < reps; ++r)
    dc_panel_spmv(s, *PK, nullptr, (const double*)s->x64, n, (const double*)s->w64, (double*)s->partK);
  HIPCHK(hipEventRecord(e2, S()));
  HIPCHK(hipEventSynchronize(e2));
  float t01 = 0.f, t12 = 0.f;
  HIPCHK(hipEventElapsedTime(&t01, e0, e1));
  HIPCHK(hipEventElapsedTime(&t12, e1, e2));
  if (ms_a) *ms_a = (double)t01 / reps;
  if (ms_k) *ms_k = (double)t12 / reps;
  if (bytes_a) *bytes_a = dc_panel_bytes(s, *PA);
  if (bytes_k) *bytes_k = dc_panel_bytes(s, *PK);
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  HIPCHK(hipEventDestroy(e2));
  // w64 and the A x cache are rebuilt by the next solve
  s->ax_valid = 0;
  return 0;
}

// Applies one of the linsys' operators to a host vector, through the same
// kernels the solve uses (panels, hybrids or row-block CSR):
//   which 0: y = K x with K = P + sigma I + A' diag(rho) A, in the CG precision
//            (x rounded to it first; y returned in FP64);
//   which 1: y = K x with the defect correction's FP64-vector operator;
//   which 2: y = A x (m entries) with the FP64-vector product of z = A x.
// The CG state is reset afterwards (the next solve starts fresh).
OSQPInt osqp_amd_linsys_apply(LinSysSolver* ls, OSQPInt which, const OSQPFloat* x, OSQPFloat* y) {
  HipPcg* s = (HipPcg*)ls;
  if (!s || which < 0 || which > 2 || !x || !y || s->polishing) return 1;
  const int n = s->n, m = s->m;
  if (which == 2 && m == 0) return 1;
  update_precond(s);
  hipLaunchKernelGGL(k_set_active, dim3(1), dim3(64), 0, S(), s->st, 1);
  upload(s->x64, x, sizeof(double) * (size_t)n);
  double* out = nullptr;
  if (which == 0) {
    if (s->bits == 32) {
      hipLaunchKernelGGL(k_cast<float>, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, S(), (long long)n,
                         (const double*)s->x64, (float*)s->p);
      if (m) op_av_rho<float>(s, (const float*)s->p, (float*)s->w);
      op_kv<float>(s, (const float*)s->p, (const float*)s->w, nullptr, (float*)s->Kp, s->parts);
      hipLaunchKernelGGL(k_widen, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, S(), n, (const float*)s->Kp,
                         s->r64);
    } else {
      if (m) op_av_rho<double>(s, (const double*)s->x64, (double*)s->w);
      op_kv<double>(s, (const double*)s->x64, (const double*)s->w, nullptr, s->r64, s->parts);
    }
    out = s->r64;
  } else if (which == 1) {
    if (m) op_av_rho<double>(s, (const double*)s->x64, s->w64);
    op_kv<double>(s, (const double*)s->x64, (const double*)s->w64, nullptr, s->r64, s->parts);
    out = s->r64;
  } else {
    double* b = (double*)dev_alloc(sizeof(double) * (size_t)(n + m));
    if (!b) return 1;
    op_finish_z(s, b, nullptr);
    download(y, b + n, sizeof(double) * (size_t)m);
    dev_free(b);
  }
  if (out) download(y, out, sizeof(double) * (size_t)n);
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, S(), s->st, s->tol_fraction);
  s->ax_valid = 0;
  s->host_solves = 0;
  HIPCHK(hipStreamSynchronize(S()));
  if (s->h_flags) memset(s->h_flags, 0, sizeof(int) * (size_t)s->flags_len);
  return osqp_amd_last_error() ? 1 : 0;
}

OSQPInt osqp_amd_linsys_bytes(LinSysSolver* ls, double* per_cg_iter, double* per_solve) {
  HipPcg* s = (HipPcg*)ls;
  const double n = s->n, m = s->m, vb = s->bits == 32 ? 4.0 : 8.0;
  const double nnzA = m ? (double)s->A->nnz_in : 0.0, nnzK = (double)s->P->nnz + nnzA;
  // one SpMV (+ its combine): panels (panel_spmv_bytes + partials read once,
  // epilogue vectors) or row-block CSR (values, 32-bit indices, row pointers,
  // x, y)
  auto av = [&](const PanelMatrix* P, double b) {
    const double epi = m * (8.0 + b);  // rho read, w written
    if (P) return panel_spmv_bytes(*P, (int)b) + (double)P->np * m * b + epi;
    const HybridOp* H = b == 4.0 ? hyb_A<float>(s) : hyb_A<double>(s);
    if (H) return hyb_bytes(*H, hyb_pv(*H, b), b) + epi;
    return nnzA * (b + 4.0) + (m + 1.0) * 4.0 + n * b + epi;
  };
  auto kv = [&](const PanelMatrix* P, double b) {
    const double epi = 3.0 * n * b;  // p (sigma p), sub or nothing, Kp written
    if (P) return panel_spmv_bytes(*P, (int)b) + (double)P->np * n * b + epi;
    const HybridOp* H = b == 4.0 ? hyb_K<float>(s) : hyb_K<double>(s);
    if (H) return hyb_bytes(*H, hyb_pv(*H, b), b) + epi;
    return nnzK * (b + 4.0) + (n + 1.0) * 4.0 + (n + m) * b + epi;
  };
  const PanelMatrix* A32 = s->bits == 32 ? panels_A<float>(s) : panels_A<double>(s);
  const PanelMatrix* K32 = s->bits == 32 ? panels_K<float>(s) : panels_K<double>(s);
  const PanelMatrix* A64 = panels_A<double>(s);
  const PanelMatrix* K64 = panels_K<double>(s);
  // CG iteration: A p, K p, update (d, p, r, Kp, dinv read; d, r, y written),
  // p update (p, y read; p written)
  double cg = (m ? av(A32, vb) : 0.0) + kv(K32, vb) + 8.0 * n * vb + 3.0 * n * vb;
  // incremental z on kZRefresh - 1 of every kZRefresh solves: A p written by
  // the combine, read by the update with u (FP32; read and written)
  const double f_incr = s->z_incr ? (kZRefresh - 1.0) / kZRefresh : 0.0;
  cg += f_incr * m * (4.0 + 4.0 + 8.0);
  // per solve: w64 from the cached A x (ax, b2, rho read; w written), FP64
  // residual, CG init (r64, dinv read; r, y, p, d written), correction (d,
  // x64 read; x64, b1 written), z = A x64 (A panels; b2 and ax written: the
  // same 16 bytes per row as av's rho read and w write)
  // the FP64-vector products (FP32 values in FP32 mode)
  auto av64 = [&](const PanelMatrix* P) {
    if (P) return dc_panel_bytes(s, *P) + (double)P->np * m * 8.0 + m * 16.0;
    return av(P, 8.0);
  };
  auto kv64 = [&](const PanelMatrix* P) {
    if (P) return dc_panel_bytes(s, *P) + (double)P->np * n * 8.0 + 3.0 * n * 8.0;
    return kv(P, 8.0);
  };
  // z: exactly (av64) on the refresh solves, else z = ax + scale u in the
  // tail (ax, u read; ax, b2 written)
  const double z = m ? (1.0 - f_incr) * av64(A64) + f_incr * m * (8.0 + 4.0 + 16.0) : 0.0;
  const double outer = 4.0 * m * 8.0 + kv64(K64) + n * (8.0 + vb) + 4.0 * n * vb + n * (vb + 8.0) +
                       2.0 * n * 8.0 + z;
  if (per_cg_iter) *per_cg_iter = cg;
  if (per_solve) *per_solve = outer;
  return 0;
}

OSQPFloat* osqp_amd_linsys_w_target(LinSysSolver* ls) {
  HipPcg* s = (HipPcg*)ls;
  if (!s || s->m <= 0 || s->polishing || !s->rho_vec || !s->ax_valid || s->ax_Aver != s->A->version) return nullptr;
  return s->w64;
}

void osqp_amd_linsys_set_w_given(LinSysSolver* ls, OSQPInt given) {
  HipPcg* s = (HipPcg*)ls;
  if (s) s->w_given = given ? 1 : 0;
}

void osqp_amd_linsys_counters(LinSysSolver* ls, OSQPInt* cg_iters, OSQPInt* solves) {
  HipPcg* s = (HipPcg*)ls;
  HIPCHK(hipMemcpyAsync(s->h_st, s->st, sizeof(PcgState), hipMemcpyDeviceToHost, S()));
  HIPCHK(hipStreamSynchronize(S()));
  if (cg_iters) *cg_iters = (OSQPInt)s->h_st->total_iters;
  if (solves) *solves = s->h_st->solves;
}

}  // extern "C"
