// pcg.hip — matrix-free Jacobi-preconditioned CG on the reduced KKT system
//     K x = (P + sigma I + A' diag(rho) A) x = b1 + A' (rho o b2)
// behind OSQP's linsys plugin vtable (include/private/types.h:243-279).
//
// Reference: algebra/cuda/lin_sys/indirect/cuda_pcg_interface.cu:32-361 and
// cuda_pcg.cu:50-285; CPU statement of K in algebra/_common/reduced_kkt.c.
//
// MI355X design:
//   * No host synchronisation inside a solve.  The loop state (r'y, ||r||,
//     alpha, beta, eps, iteration count, active flag) lives in a device
//     struct; one-block "scalar" kernels finish each reduction; every CG
//     kernel returns at once when `active` is cleared, so the host queues
//     cg_max_iter iterations and never waits.  The reference host-syncs every
//     CG iteration (cuda_pcg.cu:188,199).
//   * The tolerance schedule (cuda_pcg_interface.cu:32-64) runs on the device,
//     including the zero-iteration counter that halves the reduction factor.
//   * Two arithmetics:
//       FP64 ("plain") — the reference algorithm in double throughout.
//       FP32 ("mixed") — defect correction: the iterate x, the right-hand side
//         and the residual r = K x - rhs are FP64 (FP32-stored matrix values,
//         FP64 accumulation); the CG itself runs in FP32 on the correction
//         system K d = -r / ||r||, with FP32 lane accumulation and
//         Kahan-compensated dot products; x += ||r|| d in FP64.  Every CG
//         iteration streams 8 bytes per nonzero (FP32 value + int32 index)
//         and the solution is not limited by FP32 round-off.
//   * CSR row kernels with G lanes per row (G from the mean row length);
//     K p fuses sigma p + P p + A' w into one pass over the n rows.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hip_internal.h"

using namespace osqp_hip;

namespace {

inline hipStream_t S() { return rt().stream; }

#define GRID_STRIDE(i, n) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

struct PcgState {
  double eps;          // tolerance of the current solve (absolute, on K x = rhs)
  double eps_cmp;      // tolerance compared with the CG residual (eps / scale)
  double scale;        // defect-correction scale ||r0||_inf (1 for FP64)
  double eps_prev;     // tolerance schedule state
  double red_factor;   // tolerance schedule state
  double rhs_norm;
  double rTy;          // r'y of the current iterate
  double rnorm;        // ||r||_inf of the CG residual
  double pKp, alpha, beta;
  int iter;            // CG iterations of the current solve
  int active;          // 1 while the CG loop runs
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

// ---- w = rho o (A v), m rows.  TV: stored value type, TX: vector type -------
template <typename TV, typename TX, int G>
__global__ __launch_bounds__(kBlock) void k_Av_rho(const PcgState* st, int m, const int* __restrict__ rp,
                                                   const int* __restrict__ ci, const TV* __restrict__ val,
                                                   const TX* __restrict__ v, const double* __restrict__ rho,
                                                   double rho_sc, TX* __restrict__ w) {
  if (!st->active) return;
  const int lane = threadIdx.x & (G - 1);
  const int rpb = kBlock / G;
  for (int row = blockIdx.x * rpb + threadIdx.x / G; row < m; row += gridDim.x * rpb) {
    TX acc = 0;
    const int e = rp[row + 1];
    for (int k = rp[row] + lane; k < e; k += G) acc += (TX)val[k] * v[ci[k]];
    acc = group_sum<G>(acc);
    if (lane == 0) w[row] = (TX)rho_at(rho, rho_sc, row) * acc;
  }
}

// ---- Kv = sigma v + P v + A' w, n rows; partial sums of v'Kv ----------------
template <typename TV, typename TX, int G>
__global__ __launch_bounds__(kBlock) void k_Kv(const PcgState* st, int n, double sigma, const int* __restrict__ prp,
                                               const int* __restrict__ pci, const TV* __restrict__ pval,
                                               const int* __restrict__ trp, const int* __restrict__ tci,
                                               const TV* __restrict__ tval, const TX* __restrict__ v,
                                               const TX* __restrict__ w, TX* __restrict__ Kv, double* partials) {
  if (!st->active) return;
  __shared__ double lds[kBlock / kWave];
  const int lane = threadIdx.x & (G - 1);
  const int rpb = kBlock / G;
  const TX sg = (TX)sigma;
  Acc<TX> dot;
  for (int row = blockIdx.x * rpb + threadIdx.x / G; row < n; row += gridDim.x * rpb) {
    TX acc = 0;
    int e = prp[row + 1];
    for (int k = prp[row] + lane; k < e; k += G) acc += (TX)pval[k] * v[pci[k]];
    if (trp) {
      e = trp[row + 1];
      for (int k = trp[row] + lane; k < e; k += G) acc += (TX)tval[k] * w[tci[k]];
    }
    acc = group_sum<G>(acc);
    if (lane == 0) {
      TX vi = v[row];
      TX kv = sg * vi + acc;
      Kv[row] = kv;
      dot.add(vi * kv);
    }
  }
  if (partials) {
    double t = block_sum(dot.value(), lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
  }
}

// ---- rhs = b1 + A' (rho o b2), n rows, FP64 out; partial max|rhs| ------------
template <typename TV, int G>
__global__ __launch_bounds__(kBlock) void k_rhs(int n, const double* __restrict__ b, const int* __restrict__ trp,
                                                const int* __restrict__ tci, const TV* __restrict__ tval,
                                                const double* __restrict__ rho, double rho_sc,
                                                double* __restrict__ rhs, double* partials) {
  __shared__ double lds[kBlock / kWave];
  const int lane = threadIdx.x & (G - 1);
  const int rpb = kBlock / G;
  const double* b2 = b + n;
  double mx = 0.0;
  for (int row = blockIdx.x * rpb + threadIdx.x / G; row < n; row += gridDim.x * rpb) {
    double acc = 0.0;
    if (trp) {
      const int e = trp[row + 1];
      for (int k = trp[row] + lane; k < e; k += G) {
        int c = tci[k];
        acc += (double)tval[k] * (rho_at(rho, rho_sc, c) * b2[c]);
      }
    }
    acc = group_sum<G>(acc);
    if (lane == 0) {
      double r = b[row] + acc;
      rhs[row] = r;
      double a = fabs(r);
      mx = a > mx ? a : mx;
    }
  }
  double t = block_max(mx, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// ---- tolerance (cuda_pcg_interface.cu:32-64) + loop reset, one block ---------
__global__ void k_tolerance(PcgState* st, const double* partials, int nparts, int admm_iter, double spr, double sdr,
                            int polishing, double tol_fraction, int reduction_threshold) {
  __shared__ double lds[kBlock / kWave];
  double mx = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) mx = partials[i] > mx ? partials[i] : mx;
  double rhs_norm = block_max(mx, lds);
  if (threadIdx.x != 0) return;
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
  st->solves++;
}

// ---- FP64: r = Kx - rhs; y = M^-1 r; p = -y; partials r'y and max|r| ---------
__global__ void k_init_r(int n, const double* Kx, const double* rhs, const double* dinv, double* r, double* y,
                         double* p, double* part_rty, double* part_rn) {
  __shared__ double lds[kBlock / kWave];
  double dot = 0.0, mx = 0.0;
  GRID_STRIDE(i, n) {
    double ri = Kx[i] - rhs[i];
    double yi = dinv[i] * ri;
    r[i] = ri;
    y[i] = yi;
    p[i] = -yi;
    dot += yi * ri;
    double a = fabs(ri);
    mx = a > mx ? a : mx;
  }
  double t1 = block_sum(dot, lds);
  double t2 = block_max(mx, lds);
  if (threadIdx.x == 0) {
    part_rty[blockIdx.x] = t1;
    part_rn[blockIdx.x] = t2;
  }
}

// ---- FP32 defect correction ---------------------------------------------------
// r64 = K x64 - rhs64 (in place in Kx); partial max|r64|
__global__ void k_dc_residual(int n, double* Kx, const double* rhs, double* part_rn) {
  __shared__ double lds[kBlock / kWave];
  double mx = 0.0;
  GRID_STRIDE(i, n) {
    double ri = Kx[i] - rhs[i];
    Kx[i] = ri;
    double a = fabs(ri);
    mx = a > mx ? a : mx;
  }
  double t = block_max(mx, lds);
  if (threadIdx.x == 0) part_rn[blockIdx.x] = t;
}

// one block: scale = ||r64||; tolerance on the scaled system
__global__ void k_dc_scale(PcgState* st, const double* part_rn, int nparts) {
  __shared__ double lds[kBlock / kWave];
  double mx = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) mx = part_rn[i] > mx ? part_rn[i] : mx;
  double rn = block_max(mx, lds);
  if (threadIdx.x != 0) return;
  double s = (rn > 0.0 && isfinite(rn)) ? rn : 1.0;
  st->scale = s;
  st->eps_cmp = st->eps / s;
}

// r = r64 / scale; y = M^-1 r; p = -y; d = 0; partials r'y, max|r|
__global__ void k_dc_init_cg(const PcgState* st, int n, const double* r64, const float* dinv, float* r, float* y,
                             float* p, float* d, double* part_rty, double* part_rn) {
  __shared__ double lds[kBlock / kWave];
  const double inv = 1.0 / st->scale;
  Acc<float> dot;
  double mx = 0.0;
  GRID_STRIDE(i, n) {
    float ri = (float)(r64[i] * inv);
    float yi = dinv[i] * ri;
    r[i] = ri;
    y[i] = yi;
    p[i] = -yi;
    d[i] = 0.0f;
    dot.add(yi * ri);
    double a = fabs((double)ri);
    mx = a > mx ? a : mx;
  }
  double t1 = block_sum(dot.value(), lds);
  double t2 = block_max(mx, lds);
  if (threadIdx.x == 0) {
    part_rty[blockIdx.x] = t1;
    part_rn[blockIdx.x] = t2;
  }
}

// x64 += scale * d ; b[0:n] = x64
__global__ void k_dc_apply(const PcgState* st, int n, const float* d, double* x64, double* b) {
  const double s = st->scale;
  const int ran = st->iter > 0;
  GRID_STRIDE(i, n) {
    double xi = x64[i];
    if (ran) xi += s * (double)d[i];
    x64[i] = xi;
    b[i] = xi;
  }
}

// ---- loop head / alpha: one block -----------------------------------------------
__global__ void k_loop_head(PcgState* st, const double* part_rty, const double* part_rn, int nparts, int max_iter,
                            int first, int guard) {
  __shared__ double lds[kBlock / kWave];
  if (!st->active) return;
  double s = 0.0, mx = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    s += part_rty[i];
    mx = part_rn[i] > mx ? part_rn[i] : mx;
  }
  double rty = block_sum(s, lds);
  double rn = block_max(mx, lds);
  if (threadIdx.x != 0) return;
  if (!first) {
    st->beta = rty / st->rTy;
    st->iter++;
    st->total_iters++;
  }
  st->rTy = rty;
  st->rnorm = rn;
  int go = (rn > st->eps_cmp && st->iter < max_iter) ? 1 : 0;
  // the mixed path stops on a non-finite or non-positive r'y (FP32 breakdown)
  if (guard && !(rty > 0.0 && isfinite(rty))) go = 0;
  st->active = go;
}

__global__ void k_alpha(PcgState* st, const double* part, int nparts, int guard) {
  __shared__ double lds[kBlock / kWave];
  if (!st->active) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  double pkp = block_sum(s, lds);
  if (threadIdx.x == 0) {
    st->pKp = pkp;
    st->alpha = st->rTy / pkp;
    if (guard && !(pkp > 0.0 && isfinite(pkp))) st->active = 0;
  }
}

// x += alpha p; r += alpha Kp; y = M^-1 r; partials r'y, max|r|
template <typename V>
__global__ void k_update(const PcgState* st, int n, V* x, V* r, V* y, const V* p, const V* Kp, const V* dinv,
                         double* part_rty, double* part_rn) {
  if (!st->active) return;
  __shared__ double lds[kBlock / kWave];
  const V alpha = (V)st->alpha;
  Acc<V> dot;
  double mx = 0.0;
  GRID_STRIDE(i, n) {
    x[i] += alpha * p[i];
    V ri = r[i] + alpha * Kp[i];
    V yi = dinv[i] * ri;
    r[i] = ri;
    y[i] = yi;
    dot.add(yi * ri);
    double a = fabs((double)ri);
    mx = a > mx ? a : mx;
  }
  double t1 = block_sum(dot.value(), lds);
  double t2 = block_max(mx, lds);
  if (threadIdx.x == 0) {
    part_rty[blockIdx.x] = t1;
    part_rn[blockIdx.x] = t2;
  }
}

// p = beta p - y
template <typename V>
__global__ void k_pupdate(const PcgState* st, int n, V* p, const V* y) {
  if (!st->active) return;
  const V beta = (V)st->beta;
  GRID_STRIDE(i, n) p[i] = beta * p[i] - y[i];
}

__global__ void k_copy_x(int n, const double* x, double* b) {
  GRID_STRIDE(i, n) b[i] = x[i];
}

// b[n:n+m] = A x  (or rho_pol (A x - b2) when polishing); FP64 accumulation
template <typename TV, int G>
__global__ __launch_bounds__(kBlock) void k_finish_z(int m, int n, const int* __restrict__ rp,
                                                     const int* __restrict__ ci, const TV* __restrict__ val,
                                                     const double* __restrict__ x, double* __restrict__ b,
                                                     int polishing, double rho_pol) {
  const int lane = threadIdx.x & (G - 1);
  const int rpb = kBlock / G;
  for (int row = blockIdx.x * rpb + threadIdx.x / G; row < m; row += gridDim.x * rpb) {
    double acc = 0.0;
    const int e = rp[row + 1];
    for (int k = rp[row] + lane; k < e; k += G) acc += (double)val[k] * x[ci[k]];
    acc = group_sum<G>(acc);
    if (lane == 0) {
      double z = acc;
      if (polishing) z = (z - b[n + row]) * rho_pol;
      b[n + row] = z;
    }
  }
}

// diag preconditioner: dinv = 1 / (sigma + diag(P) + sum_j rho_j A_ji^2)
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

// ---- the solver object ----------------------------------------------------------
struct HipPcg {
  LinSysSolver base;  // must stay first: LinSysSolver* <-> HipPcg*
  int n, m, bits, polishing, precond_type;
  int max_iter, reduction_threshold;
  double tol_fraction, sigma, rho_sc;
  const OSQPMatrix *P, *A;
  const OSQPVectorf* rho_vec;
  OSQPFloat *scaled_prim_res, *scaled_dual_res;
  // FP64 vectors: x (warm start), rhs, Kx / r64, w64
  double *x64, *rhs64, *r64, *w64;
  // CG vectors in the CG precision (double for FP64, float for FP32)
  void *d, *p, *Kp, *r, *y, *w, *dinv;
  double* dtmp;       // FP64 n-vector (preconditioner diagonal)
  PcgState* st;
  double* parts;      // 3 * kMaxReduceBlocks partial slots
  PcgState* h_st;     // pinned mirror for counters
  int gA, gT, gK;     // lanes per row: A rows, A' rows, K rows
};

template <template <typename, typename, int> class K, typename TV, typename TX, typename... Args>
void launch_g2(int G, int grid, Args... args) {
  switch (G) {
    case 4: hipLaunchKernelGGL((K<TV, TX, 4>::fn), dim3(grid), dim3(kBlock), 0, S(), args...); break;
    case 8: hipLaunchKernelGGL((K<TV, TX, 8>::fn), dim3(grid), dim3(kBlock), 0, S(), args...); break;
    case 16: hipLaunchKernelGGL((K<TV, TX, 16>::fn), dim3(grid), dim3(kBlock), 0, S(), args...); break;
    case 32: hipLaunchKernelGGL((K<TV, TX, 32>::fn), dim3(grid), dim3(kBlock), 0, S(), args...); break;
    default: hipLaunchKernelGGL((K<TV, TX, 64>::fn), dim3(grid), dim3(kBlock), 0, S(), args...); break;
  }
}

template <template <typename, int> class K, typename TV, typename... Args>
void launch_g1(int G, int grid, Args... args) {
  switch (G) {
    case 4: hipLaunchKernelGGL((K<TV, 4>::fn), dim3(grid), dim3(kBlock), 0, S(), args...); break;
    case 8: hipLaunchKernelGGL((K<TV, 8>::fn), dim3(grid), dim3(kBlock), 0, S(), args...); break;
    case 16: hipLaunchKernelGGL((K<TV, 16>::fn), dim3(grid), dim3(kBlock), 0, S(), args...); break;
    case 32: hipLaunchKernelGGL((K<TV, 32>::fn), dim3(grid), dim3(kBlock), 0, S(), args...); break;
    default: hipLaunchKernelGGL((K<TV, 64>::fn), dim3(grid), dim3(kBlock), 0, S(), args...); break;
  }
}

template <typename TV, typename TX, int G> struct AvRhoK { static constexpr auto fn = k_Av_rho<TV, TX, G>; };
template <typename TV, typename TX, int G> struct KvK { static constexpr auto fn = k_Kv<TV, TX, G>; };
template <typename TV, int G> struct RhsK { static constexpr auto fn = k_rhs<TV, G>; };
template <typename TV, int G> struct FinishK { static constexpr auto fn = k_finish_z<TV, G>; };

inline int rows_grid(int rows, int G, int cap) {
  const int rpb = kBlock / G;
  long long g = (rows + rpb - 1) / rpb;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

template <typename TV>
struct Views {
  const int *arp, *aci, *trp, *tci, *prp, *pci;
  const TV *aval, *tval, *pval;
};

template <typename TV>
Views<TV> views(const HipPcg* s);

template <>
Views<double> views<double>(const HipPcg* s) {
  Views<double> v;
  v.arp = s->A->rp; v.aci = s->A->ci; v.aval = s->A->val;
  v.trp = s->m ? s->A->trp : nullptr; v.tci = s->A->tci; v.tval = s->A->tval;
  v.prp = s->P->rp; v.pci = s->P->ci; v.pval = s->P->val;
  return v;
}

template <>
Views<float> views<float>(const HipPcg* s) {
  Views<float> v;
  v.arp = s->A->rp; v.aci = s->A->ci; v.aval = s->A->val32;
  v.trp = s->m ? s->A->trp : nullptr; v.tci = s->A->tci; v.tval = s->A->tval32;
  v.prp = s->P->rp; v.pci = s->P->ci; v.pval = s->P->val32;
  return v;
}

void refresh_fp32(const OSQPMatrix* Mc) {
  OSQPMatrix* M = (OSQPMatrix*)Mc;
  if (!M->fp32_dirty) return;
  if (!M->val32) M->val32 = (float*)dev_alloc(sizeof(float) * (size_t)(M->nnz > 0 ? M->nnz : 1));
  if (M->nnz > 0)
    hipLaunchKernelGGL(k_cast<float>, dim3(grid_for(M->nnz, kBlock, 16384)), dim3(kBlock), 0, S(), (long long)M->nnz,
                       (const double*)M->val, M->val32);
  if (!M->is_triu) {
    if (!M->tval32) M->tval32 = (float*)dev_alloc(sizeof(float) * (size_t)(M->nnz_in > 0 ? M->nnz_in : 1));
    if (M->nnz_in > 0)
      hipLaunchKernelGGL(k_cast<float>, dim3(grid_for(M->nnz_in, kBlock, 16384)), dim3(kBlock), 0, S(),
                         (long long)M->nnz_in, (const double*)M->tval, M->tval32);
  }
  M->fp32_dirty = 0;
}

template <int G> struct PrecK { static constexpr auto fn = k_precond<G>; };

void update_precond(HipPcg* s) {
  if (s->bits == 32) {
    refresh_fp32(s->P);
    refresh_fp32(s->A);
  }
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

// CG loop shared by both arithmetics: x is the CG iterate (x64 or d)
template <typename V>
void cg_loop(HipPcg* s, V* x, int gA, int gK, int gE, int guard) {
  const int n = s->n, m = s->m;
  Views<V> mv = views<V>(s);
  V* p = (V*)s->p;
  V* Kp = (V*)s->Kp;
  V* r = (V*)s->r;
  V* y = (V*)s->y;
  V* w = (V*)s->w;
  const V* dinv = (const V*)s->dinv;
  const double* rho = s->rho_vec ? s->rho_vec->d : nullptr;
  double* part0 = s->parts;
  double* part1 = s->parts + kMaxReduceBlocks;
  double* part2 = s->parts + 2 * kMaxReduceBlocks;
  for (int it = 0; it < s->max_iter; ++it) {
    if (m)
      launch_g2<AvRhoK, V, V>(s->gA, gA, (const PcgState*)s->st, m, mv.arp, mv.aci, mv.aval, (const V*)p, rho,
                              s->rho_sc, w);
    launch_g2<KvK, V, V>(s->gK, gK, (const PcgState*)s->st, n, s->sigma, mv.prp, mv.pci, mv.pval,
                         m ? mv.trp : nullptr, mv.tci, mv.tval, (const V*)p, (const V*)w, Kp, part2);
    hipLaunchKernelGGL(k_alpha, dim3(1), dim3(kBlock), 0, S(), s->st, (const double*)part2, gK, guard);
    hipLaunchKernelGGL(k_update<V>, dim3(gE), dim3(kBlock), 0, S(), (const PcgState*)s->st, n, x, r, y,
                       (const V*)p, (const V*)Kp, dinv, part0, part1);
    hipLaunchKernelGGL(k_loop_head, dim3(1), dim3(kBlock), 0, S(), s->st, (const double*)part0,
                       (const double*)part1, gE, s->max_iter, 0, guard);
    hipLaunchKernelGGL(k_pupdate<V>, dim3(gE), dim3(kBlock), 0, S(), (const PcgState*)s->st, n, p, (const V*)y);
  }
}

void solve_fp64(HipPcg* s, OSQPVectorf* b, int admm_iter) {
  const int n = s->n, m = s->m;
  Views<double> mv = views<double>(s);
  const double* rho = s->rho_vec ? s->rho_vec->d : nullptr;
  double* part0 = s->parts;
  double* part1 = s->parts + kMaxReduceBlocks;
  double* part2 = s->parts + 2 * kMaxReduceBlocks;
  const int gK = rows_grid(n, s->gK, kMaxReduceBlocks);
  const int gA = rows_grid(m, s->gA, 65535 * 4);
  const int gE = grid_for(n, kBlock, kMaxReduceBlocks);
  double* r = (double*)s->r;
  double* y = (double*)s->y;
  double* p = (double*)s->p;
  launch_g1<RhsK, double>(s->gK, gK, n, (const double*)b->d, mv.trp, mv.tci, mv.tval, rho, s->rho_sc, s->rhs64,
                          part0);
  hipLaunchKernelGGL(k_tolerance, dim3(1), dim3(kBlock), 0, S(), s->st, (const double*)part0, gK, admm_iter,
                     *s->scaled_prim_res, *s->scaled_dual_res, s->polishing, s->tol_fraction,
                     s->reduction_threshold);
  // r = K x - rhs (x warm-started from the previous solve)
  if (m)
    launch_g2<AvRhoK, double, double>(s->gA, gA, (const PcgState*)s->st, m, mv.arp, mv.aci, mv.aval,
                                      (const double*)s->x64, rho, s->rho_sc, s->w64);
  launch_g2<KvK, double, double>(s->gK, gK, (const PcgState*)s->st, n, s->sigma, mv.prp, mv.pci, mv.pval,
                                 m ? mv.trp : nullptr, mv.tci, mv.tval, (const double*)s->x64,
                                 (const double*)s->w64, s->r64, (double*)nullptr);
  hipLaunchKernelGGL(k_init_r, dim3(gE), dim3(kBlock), 0, S(), n, (const double*)s->r64, (const double*)s->rhs64,
                     (const double*)s->dinv, r, y, p, part0, part1);
  hipLaunchKernelGGL(k_loop_head, dim3(1), dim3(kBlock), 0, S(), s->st, (const double*)part0, (const double*)part1,
                     gE, s->max_iter, 1, 0);
  cg_loop<double>(s, s->x64, gA, gK, gE, 0);
  hipLaunchKernelGGL(k_copy_x, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, S(), n, (const double*)s->x64, b->d);
  if (m)
    launch_g1<FinishK, double>(s->gA, gA, m, n, mv.arp, mv.aci, mv.aval, (const double*)s->x64, b->d, s->polishing,
                               s->rho_sc);
  LAUNCHCHK();
  (void)part2;
}

void solve_fp32(HipPcg* s, OSQPVectorf* b, int admm_iter) {
  const int n = s->n, m = s->m;
  Views<float> mv = views<float>(s);
  const double* rho = s->rho_vec ? s->rho_vec->d : nullptr;
  double* part0 = s->parts;
  double* part1 = s->parts + kMaxReduceBlocks;
  const int gK = rows_grid(n, s->gK, kMaxReduceBlocks);
  const int gA = rows_grid(m, s->gA, 65535 * 4);
  const int gE = grid_for(n, kBlock, kMaxReduceBlocks);
  launch_g1<RhsK, float>(s->gK, gK, n, (const double*)b->d, mv.trp, mv.tci, mv.tval, rho, s->rho_sc, s->rhs64,
                         part0);
  hipLaunchKernelGGL(k_tolerance, dim3(1), dim3(kBlock), 0, S(), s->st, (const double*)part0, gK, admm_iter,
                     *s->scaled_prim_res, *s->scaled_dual_res, s->polishing, s->tol_fraction,
                     s->reduction_threshold);
  // r64 = K x64 - rhs64 with FP32 values and FP64 accumulation
  if (m)
    launch_g2<AvRhoK, float, double>(s->gA, gA, (const PcgState*)s->st, m, mv.arp, mv.aci, mv.aval,
                                     (const double*)s->x64, rho, s->rho_sc, s->w64);
  launch_g2<KvK, float, double>(s->gK, gK, (const PcgState*)s->st, n, s->sigma, mv.prp, mv.pci, mv.pval,
                                m ? mv.trp : nullptr, mv.tci, mv.tval, (const double*)s->x64,
                                (const double*)s->w64, s->r64, (double*)nullptr);
  hipLaunchKernelGGL(k_dc_residual, dim3(gE), dim3(kBlock), 0, S(), n, s->r64, (const double*)s->rhs64, part1);
  hipLaunchKernelGGL(k_dc_scale, dim3(1), dim3(kBlock), 0, S(), s->st, (const double*)part1, gE);
  hipLaunchKernelGGL(k_dc_init_cg, dim3(gE), dim3(kBlock), 0, S(), (const PcgState*)s->st, n, (const double*)s->r64,
                     (const float*)s->dinv, (float*)s->r, (float*)s->y, (float*)s->p, (float*)s->d, part0, part1);
  hipLaunchKernelGGL(k_loop_head, dim3(1), dim3(kBlock), 0, S(), s->st, (const double*)part0, (const double*)part1,
                     gE, s->max_iter, 1, 1);
  cg_loop<float>(s, (float*)s->d, gA, gK, gE, 1);
  hipLaunchKernelGGL(k_dc_apply, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, S(), (const PcgState*)s->st, n,
                     (const float*)s->d, s->x64, b->d);
  if (m)
    launch_g1<FinishK, float>(s->gA, gA, m, n, mv.arp, mv.aci, mv.aval, (const double*)s->x64, b->d, s->polishing,
                              s->rho_sc);
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
  if (s->bits == 32) {
    if (s->A->fp32_dirty || s->P->fp32_dirty) update_precond(s);
    solve_fp32(s, b, admm_iter);
  } else {
    solve_fp64(s, b, admm_iter);
  }
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
}

void pcg_free(LinSysSolver* self) {
  HipPcg* s = (HipPcg*)self;
  if (!s) return;
  HIPCHK(hipStreamSynchronize(S()));
  dev_free(s->x64); dev_free(s->rhs64); dev_free(s->r64); dev_free(s->w64);
  dev_free(s->d); dev_free(s->p); dev_free(s->Kp); dev_free(s->r); dev_free(s->y);
  dev_free(s->w); dev_free(s->dinv); dev_free(s->dtmp);
  dev_free(s->st); dev_free(s->parts);
  if (s->h_st) (void)hipHostFree(s->h_st);
  free(s);
}

OSQPInt pcg_update_matrices(LinSysSolver* self, const OSQPMatrix*, const OSQPInt*, OSQPInt, const OSQPMatrix*,
                            const OSQPInt*, OSQPInt) {
  update_precond((HipPcg*)self);
  return 0;
}

OSQPInt pcg_update_rho(LinSysSolver* self, const OSQPVectorf*, OSQPFloat rho_sc) {
  HipPcg* s = (HipPcg*)self;
  s->rho_sc = rho_sc;
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
  s->d = dev_alloc(vb * n);
  s->p = dev_alloc(vb * n);
  s->Kp = dev_alloc(vb * n);
  s->r = dev_alloc(vb * n);
  s->y = dev_alloc(vb * n);
  s->w = dev_alloc(vb * m);
  s->dinv = dev_alloc(vb * n);
  s->dtmp = (double*)dev_alloc(sizeof(double) * n);
  s->st = (PcgState*)dev_alloc(sizeof(PcgState));
  s->parts = (double*)dev_alloc(sizeof(double) * 3 * kMaxReduceBlocks);
  if (!s->x64 || !s->rhs64 || !s->r64 || !s->w64 || !s->d || !s->p || !s->Kp || !s->r || !s->y || !s->w ||
      !s->dinv || !s->dtmp || !s->st || !s->parts) {
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
  update_precond(s);
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
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, S(), s->st, s->tol_fraction);
  // activate the loop flag so the kernels run their full body
  PcgState h;
  memset(&h, 0, sizeof(h));
  h.active = 1;
  h.red_factor = s->tol_fraction;
  h.eps_prev = 1.0;
  h.scale = 1.0;
  HIPCHK(hipMemcpyAsync(s->st, &h, sizeof(h), hipMemcpyHostToDevice, S()));
  HIPCHK(hipStreamSynchronize(S()));
  const int gK = rows_grid(n, s->gK, kMaxReduceBlocks);
  const int gA = rows_grid(m, s->gA, 65535 * 4);
  const double* rho = s->rho_vec ? s->rho_vec->d : nullptr;
  double* part2 = s->parts + 2 * kMaxReduceBlocks;
  HIPCHK(hipEventRecord(e0, S()));
  for (int r = 0; r < reps; ++r) {
    if (s->bits == 32) {
      Views<float> mv = views<float>(s);
      if (m) launch_g2<AvRhoK, float, float>(s->gA, gA, (const PcgState*)s->st, m, mv.arp, mv.aci, mv.aval,
                                             (const float*)s->p, rho, s->rho_sc, (float*)s->w);
    } else {
      Views<double> mv = views<double>(s);
      if (m) launch_g2<AvRhoK, double, double>(s->gA, gA, (const PcgState*)s->st, m, mv.arp, mv.aci, mv.aval,
                                               (const double*)s->p, rho, s->rho_sc, (double*)s->w);
    }
  }
  HIPCHK(hipEventRecord(e1, S()));
  for (int r = 0; r < reps; ++r) {
    if (s->bits == 32) {
      Views<float> mv = views<float>(s);
      launch_g2<KvK, float, float>(s->gK, gK, (const PcgState*)s->st, n, s->sigma, mv.prp, mv.pci, mv.pval,
                                   m ? mv.trp : nullptr, mv.tci, mv.tval, (const float*)s->p, (const float*)s->w,
                                   (float*)s->Kp, part2);
    } else {
      Views<double> mv = views<double>(s);
      launch_g2<KvK, double, double>(s->gK, gK, (const PcgState*)s->st, n, s->sigma, mv.prp, mv.pci, mv.pval,
                                     m ? mv.trp : nullptr, mv.tci, mv.tval, (const double*)s->p,
                                     (const double*)s->w, (double*)s->Kp, part2);
    }
  }
  HIPCHK(hipEventRecord(e2, S()));
  HIPCHK(hipEventSynchronize(e2));
  float t01 = 0.f, t12 = 0.f;
  HIPCHK(hipEventElapsedTime(&t01, e0, e1));
  HIPCHK(hipEventElapsedTime(&t12, e1, e2));
  if (ms_av) *ms_av = (double)t01 / reps;
  if (ms_kv) *ms_kv = (double)t12 / reps;
  const double nnzA = (double)s->A->nnz_in, nnzP = (double)s->P->nnz;
  if (bytes_av) *bytes_av = nnzA * (vb + 4.0) + (m + 1.0) * 4.0 + n * vb + m * 8.0 + m * vb;
  if (bytes_kv)
    *bytes_kv = nnzP * (vb + 4.0) + (n + 1.0) * 4.0 + nnzA * (vb + 4.0) + (n + 1.0) * 4.0 + m * vb + 2.0 * n * vb;
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  HIPCHK(hipEventDestroy(e2));
  // leave the state as a fresh solver's
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, S(), s->st, s->tol_fraction);
  HIPCHK(hipStreamSynchronize(S()));
  return 0;
}

void osqp_amd_linsys_counters(LinSysSolver* ls, OSQPInt* cg_iters, OSQPInt* solves) {
  HipPcg* s = (HipPcg*)ls;
  HIPCHK(hipMemcpyAsync(s->h_st, s->st, sizeof(PcgState), hipMemcpyDeviceToHost, S()));
  HIPCHK(hipStreamSynchronize(S()));
  if (cg_iters) *cg_iters = (OSQPInt)s->h_st->total_iters;
  if (solves) *solves = s->h_st->solves;
}

}  // extern "C"
