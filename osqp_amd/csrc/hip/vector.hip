// vector.hip — OSQPVectorf/OSQPVectori on HBM.
//
// Semantics follow the builtin backend (algebra/builtin/vector.c:9-953),
// including its operation order, so that every elementwise result is the
// same correctly rounded IEEE value (compiled with -ffp-contract=off).
// Reductions are deterministic two-pass trees (per-block partials, then one
// block), never atomics; max-type reductions are therefore bit-exact against
// the CPU and sums differ only by summation order.
#include <math.h>
#include <time.h>

#include <mutex>

#include "hip_internal.h"

using namespace osqp_hip;

namespace {

// ---- elementwise kernels ---------------------------------------------------
// c_max / c_min of the reference (glob_opts.h): ternaries, not fmax/fmin,
// so signed zeros and ties resolve exactly as on the CPU.
__device__ inline double cmax(double a, double b) { return a > b ? a : b; }
__device__ inline double cmin(double a, double b) { return a < b ? a : b; }

#define GRID_STRIDE(i, n) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

__global__ void k_set(double* a, double sc, int n) { GRID_STRIDE(i, n) a[i] = sc; }
__global__ void k_set_int(int* a, int sc, int n) { GRID_STRIDE(i, n) a[i] = sc; }

__global__ void k_set_cond(double* a, const int* t, double neg, double zero, double pos, int n) {
  GRID_STRIDE(i, n) {
    int v = t[i];
    a[i] = (v == 0) ? zero : ((v > 0) ? pos : neg);
  }
}

__global__ void k_round_to_zero(double* a, double tol, int n) {
  GRID_STRIDE(i, n) if (fabs(a[i]) < tol) a[i] = 0.0;
}

__global__ void k_mult_sc(double* a, double sc, int n) { GRID_STRIDE(i, n) a[i] *= sc; }

__global__ void k_plus(double* x, const double* a, const double* b, int n) { GRID_STRIDE(i, n) x[i] = a[i] + b[i]; }
__global__ void k_minus(double* x, const double* a, const double* b, int n) { GRID_STRIDE(i, n) x[i] = a[i] - b[i]; }

__global__ void k_add_scaled(double* x, double sa, const double* a, double sb, const double* b, int n) {
  GRID_STRIDE(i, n) x[i] = sa * a[i] + sb * b[i];
}

// vector.c:439-466: the increment form x += (sb*b + sc*c) rounds differently
// from sa*a + sb*b + sc*c, so both forms are kept.
__global__ void k_add_scaled3(double* x, double sa, const double* a, double sb, const double* b, double sc,
                              const double* c, int incr, int n) {
  if (incr) {
    GRID_STRIDE(i, n) x[i] += sb * b[i] + sc * c[i];
  } else {
    GRID_STRIDE(i, n) x[i] = sa * a[i] + sb * b[i] + sc * c[i];
  }
}

__global__ void k_ew_prod(double* c, const double* a, const double* b, int n) { GRID_STRIDE(i, n) c[i] = a[i] * b[i]; }

__global__ void k_bound(double* x, const double* z, const double* l, const double* u, int n) {
  GRID_STRIDE(i, n) x[i] = cmin(cmax(z[i], l[i]), u[i]);
}

__global__ void k_polar_reccone(double* y, const double* l, const double* u, double infval, int n) {
  GRID_STRIDE(i, n) {
    if (u[i] > +infval) {
      if (l[i] < -infval) y[i] = 0.0;
      else y[i] = cmin(y[i], 0.0);
    } else if (l[i] < -infval) {
      y[i] = cmax(y[i], 0.0);
    }
  }
}

__global__ void k_reciprocal(double* b, const double* a, int n) { GRID_STRIDE(i, n) b[i] = 1.0 / a[i]; }
__global__ void k_sqrt(double* a, int n) { GRID_STRIDE(i, n) a[i] = sqrt(a[i]); }
__global__ void k_max_vec(double* c, const double* a, const double* b, int n) {
  GRID_STRIDE(i, n) c[i] = a[i] > b[i] ? a[i] : b[i];
}
__global__ void k_min_vec(double* c, const double* a, const double* b, int n) {
  GRID_STRIDE(i, n) c[i] = a[i] < b[i] ? a[i] : b[i];
}
__global__ void k_if_lt(double* x, const double* z, double t, double v, int n) {
  GRID_STRIDE(i, n) x[i] = z[i] < t ? v : z[i];
}
__global__ void k_if_gt(double* x, const double* z, double t, double v, int n) {
  GRID_STRIDE(i, n) x[i] = z[i] > t ? v : z[i];
}

// ---- reductions --------------------------------------------------------------
enum RedOp : int {
  R_NORM_INF = 0,
  R_SCALED_NORM_INF,
  R_NORM_INF_DIFF,
  R_NORM_1,
  R_DOT,
  R_DOT_POS,
  R_DOT_NEG,
  R_ANY_GT,        // max over i of (a[i] > b[i])       -> all_leq = !result
  R_RECCONE_BAD,   // max over i of reccone violation  -> in_reccone = !result
  R_EQ_BAD,        // max over i of |a-b| > tol        -> is_eq = !result
};

struct RedArgs {
  const double* a;
  const double* b;
  const double* c;
  const double* d;
  double p0, p1;
};

__device__ inline bool is_max_op(int op) { return op != R_NORM_1 && op != R_DOT && op != R_DOT_POS && op != R_DOT_NEG; }

__device__ inline double red_elem(int op, const RedArgs& r, long long i) {
  switch (op) {
    case R_NORM_INF: return fabs(r.a[i]);
    case R_SCALED_NORM_INF: return fabs(r.a[i] * r.b[i]);
    case R_NORM_INF_DIFF: return fabs(r.a[i] - r.b[i]);
    case R_NORM_1: return fabs(r.a[i]);
    case R_DOT: return r.a[i] * r.b[i];
    case R_DOT_POS: return r.a[i] * cmax(r.b[i], 0.0);
    case R_DOT_NEG: return r.a[i] * cmin(r.b[i], 0.0);
    case R_ANY_GT: return r.a[i] > r.b[i] ? 1.0 : 0.0;
    case R_RECCONE_BAD: {
      // vector.c:704-727 with y=a, l=b, u=c, infval=p0, tol=p1
      double y = r.a[i], l = r.b[i], u = r.c[i];
      bool bad = ((u < +r.p0) && (y > +r.p1)) || ((l > -r.p0) && (y < -r.p1));
      return bad ? 1.0 : 0.0;
    }
    case R_EQ_BAD: return fabs(r.a[i] - r.b[i]) > r.p0 ? 1.0 : 0.0;
  }
  return 0.0;
}

// The last block of a reducing kernel to finish (counted on `ticket`) adds
// the per-block partials in index order -- the same order whatever the block
// timing, so results are deterministic -- writes the result to host-visible
// memory (out[0], then the sequence number `seq` in out[1]) and re-arms the
// ticket for the next reduction on the stream.
// With to_device, the result goes to out[0] in device memory alone (no
// sequence number, no system-scope fence): the caller orders on its stream.
template <bool to_device = false>
__device__ void finish_reduce_n(bool mx, double t, double* partials, unsigned* ticket, double* out, double seq,
                                double* lds, int nblocks) {
  __shared__ int last;
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = t;
    __threadfence();  // the partial is visible device-wide before the ticket
    last = atomicAdd(ticket, 1u) == (unsigned)nblocks - 1;
    if (last) __threadfence();  // every other block's partial is visible here
  }
  __syncthreads();
  if (!last) return;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += blockDim.x) {
    const double v = partials[i];
    if (mx) acc = v > acc ? v : acc;
    else acc += v;
  }
  const double r = mx ? block_max(acc, lds) : block_sum(acc, lds);
  if (threadIdx.x == 0) {
    *ticket = 0;
    out[0] = r;
    if (!to_device) {
      __threadfence_system();  // the result reaches the host before its sequence number
      ((volatile double*)out)[1] = seq;
    }
  }
}

__device__ void finish_reduce(bool mx, double t, double* partials, unsigned* ticket, double* out, double seq,
                              double* lds) {
  finish_reduce_n(mx, t, partials, ticket, out, seq, lds, (int)gridDim.x);
}

// A thread takes elements i0, i0 + S, i0 + 2S, ... (S = the grid's thread
// count) in order, with the loads of eight of them in flight at a time.
constexpr int kRedUnroll = 8;
__global__ void k_reduce(int op, RedArgs r, int n, double* partials, unsigned* ticket, double* out, double seq) {
  __shared__ double lds[kBlock / kWave];
  const bool mx = is_max_op(op);
  const long long S = (long long)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += kRedUnroll * S) {
    double v[kRedUnroll];
#pragma unroll
    for (int u = 0; u < kRedUnroll; ++u) {
      const long long i = i0 + u * S;
      v[u] = i < n ? red_elem(op, r, i) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kRedUnroll; ++u) {
      if (i0 + u * S < n) {
        if (mx) acc = v[u] > acc ? v[u] : acc;
        else acc += v[u];
      }
    }
  }
  const double t = mx ? block_max(acc, lds) : block_sum(acc, lds);
  finish_reduce(mx, t, partials, ticket, out, seq, lds);
}

// Several reductions in one launch: row k of the grid (blockIdx.y) is
// reduction k, run exactly as k_reduce would run it alone on a grid of
// t[k].g blocks (the same elements per lane, the same order, so the same
// bits), with its own partials, ticket and result slot k + 1.
constexpr int kMultiMax = 16;
struct RedTask {
  int op, n, g, pad;
  RedArgs r;
};
struct RedBatch {
  RedTask t[kMultiMax];
};

// With to_device, reduction k's result goes to out[k] in device memory.
template <bool to_device>
__global__ void k_reduce_multi(RedBatch B, double* partials, unsigned* tickets, double* out, double seq0) {
  const int k = blockIdx.y;
  const int op = B.t[k].op, n = B.t[k].n, g = B.t[k].g;
  const RedArgs r = B.t[k].r;
  if ((int)blockIdx.x >= g) return;
  __shared__ double lds[kBlock / kWave];
  const bool mx = is_max_op(op);
  const long long S = (long long)g * blockDim.x;
  double acc = 0.0;
  for (long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += kRedUnroll * S) {
    double v[kRedUnroll];
#pragma unroll
    for (int u = 0; u < kRedUnroll; ++u) {
      const long long i = i0 + u * S;
      v[u] = i < n ? red_elem(op, r, i) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kRedUnroll; ++u) {
      if (i0 + u * S < n) {
        if (mx) acc = v[u] > acc ? v[u] : acc;
        else acc += v[u];
      }
    }
  }
  const double t = mx ? block_max(acc, lds) : block_sum(acc, lds);
  finish_reduce_n<to_device>(mx, t, partials + (long long)k * 512, tickets + 1 + k,
                             to_device ? out + k : out + 2 * (k + 1), seq0 + k, lds, g);
}

// The host waits for reduction `seq` by watching its sequence slot (a stream
// synchronisation sleeps and wakes tens of microseconds late, and the
// termination checks run a dozen reductions back to back).  After half a
// second it synchronises the stream instead, which also reports a failed
// kernel.
double wait_result(Runtime& R, double seq) {
  volatile double* h = R.h_scalars;
  struct timespec t0, t;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (long k = 1;; ++k) {
    if (h[1] == seq) return h[0];
    if ((k & 1023) == 0) {
      clock_gettime(CLOCK_MONOTONIC, &t);
      if ((t.tv_sec - t0.tv_sec) + 1e-9 * (t.tv_nsec - t0.tv_nsec) > 0.5) break;
    }
  }
  HIPCHK(hipStreamSynchronize(R.stream));
  return h[0];
}

// Queued reductions are recorded on the host and run together, in one
// k_reduce_multi launch, when they are waited for: each reads its vectors as
// they are at that moment, so a caller queues a reduction only once its
// inputs are final.  Reduction k writes [value, seq] to slot k + 1 of the
// host-coherent scalars (slot 0 is the synchronous path's).
inline int red_grid(int n) {
  // about one element per lane per unrolled round: few blocks, so few
  // ticket atomics on the one counter
  const long long gg = (n + (long long)kBlock * kRedUnroll - 1) / ((long long)kBlock * kRedUnroll);
  return (int)(gg < 1 ? 1 : (gg > 512 ? 512 : gg));
}

struct RedQueue {
  RedBatch b;
  double* dst[kMultiMax];
  int count = 0;
};
RedQueue g_redq;

void red_drain(Runtime& R) {
  RedQueue& Q = g_redq;
  const int k = Q.count;
  if (k == 0) return;
  Q.count = 0;
  int gmax = 1;
  for (int j = 0; j < k; ++j) gmax = Q.b.t[j].g > gmax ? Q.b.t[j].g : gmax;
  const double seq0 = R.red_seq + 1;
  R.red_seq += k;
  hipLaunchKernelGGL(k_reduce_multi<false>, dim3(gmax, k), dim3(kBlock), 0, R.stream, Q.b, R.d_partials,
                     R.d_ticket, R.dh_scalars, seq0);
  if (LAUNCHCHK()) {
    for (int j = 0; j < k; ++j) *Q.dst[j] = 0.0;
    return;
  }
  volatile double* h = R.h_scalars;
  struct timespec t0, t;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int seen = 0;
  for (long it = 1; seen < k; ++it) {
    while (seen < k && h[2 * (seen + 1) + 1] == seq0 + seen) ++seen;
    if (seen < k && (it & 1023) == 0) {
      clock_gettime(CLOCK_MONOTONIC, &t);
      if ((t.tv_sec - t0.tv_sec) + 1e-9 * (t.tv_nsec - t0.tv_nsec) > 0.5) break;
    }
  }
  if (seen < k) HIPCHK(hipStreamSynchronize(R.stream));
  for (int j = 0; j < k; ++j) *Q.dst[j] = h[2 * (j + 1)];
}

void red_queue(int op, const RedArgs& r, int n, double* dst) {
  Runtime& R = rt();
  if (n <= 0) {
    *dst = 0.0;
    return;
  }
  RedQueue& Q = g_redq;
  if (Q.count == kMultiMax) red_drain(R);
  RedTask& T = Q.b.t[Q.count];
  T.op = op;
  T.n = n;
  T.g = red_grid(n);
  T.pad = 0;
  T.r = r;
  Q.dst[Q.count++] = dst;
}

// one kernel; the host reads the result from host-coherent memory
double reduce(int op, const RedArgs& r, int n) {
  if (n <= 0) return 0.0;
  Runtime& R = rt();
  red_drain(R);
  const int g = red_grid(n);
  const double seq = ++R.red_seq;
  hipLaunchKernelGGL(k_reduce, dim3(g), dim3(kBlock), 0, R.stream, op, r, n, R.d_partials, R.d_ticket,
                     R.dh_scalars, seq);
  if (LAUNCHCHK()) return 0.0;
  return wait_result(R, seq);
}

// ew_bounds_type: classify constraints (vector.c:885-921; loose checked
// before equality, as the builtin does) and report whether any type changed.
__global__ void k_bounds_type(int* iseq, const double* l, const double* u, double tol, double infval, int n,
                              double* partials, unsigned* ticket, double* out, double seq) {
  __shared__ double lds[kBlock / kWave];
  double changed = 0.0;
  GRID_STRIDE(i, n) {
    int old = iseq[i], t;
    if ((l[i] < -infval) && (u[i] > infval)) t = -1;
    else if (u[i] - l[i] < tol) t = 1;
    else t = 0;
    iseq[i] = t;
    if (t != old) changed = 1.0;
  }
  const double tt = block_max(changed, lds);
  finish_reduce(true, tt, partials, ticket, out, seq, lds);
}

inline int egrid(int n) { return grid_for(n, kBlock, 8192); }

inline hipStream_t S() { return rt().stream; }

}  // namespace

// ---- fused ADMM steps (host core, rho vector case) ---------------------------
// One pass over [0, n) and [n, n + m) each; the element arithmetic is that of
// the unfused algebra calls, in the same order, so results are bit-identical.
//   compute_rhs (auxil.c:136-158): xt = sigma x_prev + (-1) q;
//                                  zt = (-1)(rhoinv o y) + 1 z_prev
__global__ void k_admm_rhs(int n, int m, double sigma, double* xt, const double* xp, const double* q, double* zt,
                           const double* rhoinv, const double* y, const double* zp) {
  GRID_STRIDE(k, (long long)n + m) {
    if (k < n) {
      xt[k] = sigma * xp[k] + (-1.0) * q[k];
    } else {
      const long long i = k - n;
      const double v = rhoinv[i] * y[i];
      zt[i] = (-1.0) * v + 1.0 * zp[i];
    }
  }
}
//   update_x (auxil.c:172-184), update_z (:186-208), update_y (:210-229);
//   with `next`, also the next iteration's compute_rhs from the new x, z, y
//   (written over xt / zt once read; the arithmetic of k_admm_rhs), and with
//   wn the PCG's w = rho o (A x - b2) for that right-hand side: A x of the
//   last solve is the zt read here (the solve returns z~ = A x), so this is
//   the arithmetic of the PCG's k_tol_w on the same values
__global__ void k_admm_update(int n, int m, double a, double* x, double* dx, double* xt, const double* xp,
                              double* z, double* y, double* dy, double* zt, const double* zp,
                              const double* rho, const double* rhoinv, const double* l, const double* u, int next,
                              double sigma, const double* q, double* wn) {
  GRID_STRIDE(k, (long long)n + m) {
    if (k < n) {
      const double v = a * xt[k] + (1.0 - a) * xp[k];
      x[k] = v;
      dx[k] = v - xp[k];
      if (next) xt[k] = sigma * v + (-1.0) * q[k];
    } else {
      const long long i = k - n;
      const double ztk = zt[i];
      double v = rhoinv[i] * y[i];
      v += a * ztk + (1.0 - a) * zp[i];  // add_scaled3's increment form (x == a, sca == 1)
      v = cmin(cmax(v, l[i]), u[i]);
      z[i] = v;
      double d = a * ztk + (1.0 - a) * zp[i] + (-1.0) * v;
      d = d * rho[i];
      dy[i] = d;
      const double yn = y[i] + d;
      y[i] = yn;
      if (next) {
        const double w = rhoinv[i] * yn;
        const double ztn = (-1.0) * w + 1.0 * v;
        zt[i] = ztn;
        if (wn) wn[i] = rho[i] * (ztk - ztn);
      }
    }
  }
}

extern "C" {

// ---- creation / views / copies ------------------------------------------------
OSQPVectorf* OSQPVectorf_malloc(OSQPInt length) {
  OSQPVectorf* v = (OSQPVectorf*)calloc(1, sizeof(OSQPVectorf));
  if (!v) return nullptr;
  v->length = length;
  v->is_view = 0;
  v->d = length > 0 ? (double*)dev_alloc(sizeof(double) * (size_t)length) : nullptr;
  if (length > 0 && !v->d) {
    free(v);
    return nullptr;
  }
  return v;
}

OSQPVectori* OSQPVectori_malloc(OSQPInt length) {
  OSQPVectori* v = (OSQPVectori*)calloc(1, sizeof(OSQPVectori));
  if (!v) return nullptr;
  v->length = length;
  v->d = length > 0 ? (int*)dev_alloc(sizeof(int) * (size_t)length) : nullptr;
  if (length > 0 && !v->d) {
    free(v);
    return nullptr;
  }
  return v;
}

OSQPVectorf* OSQPVectorf_calloc(OSQPInt length) {
  OSQPVectorf* v = OSQPVectorf_malloc(length);
  if (v && length > 0) hipLaunchKernelGGL(k_set, dim3(egrid(length)), dim3(kBlock), 0, S(), v->d, 0.0, length);
  return v;
}

OSQPVectori* OSQPVectori_calloc(OSQPInt length) {
  OSQPVectori* v = OSQPVectori_malloc(length);
  if (v && length > 0) hipLaunchKernelGGL(k_set_int, dim3(egrid(length)), dim3(kBlock), 0, S(), v->d, 0, length);
  return v;
}

OSQPVectorf* OSQPVectorf_new(const OSQPFloat* a, OSQPInt length) {
  OSQPVectorf* v = OSQPVectorf_malloc(length);
  if (v && length > 0) OSQPVectorf_from_raw(v, a);
  return v;
}

OSQPVectori* OSQPVectori_new(const OSQPInt* a, OSQPInt length) {
  OSQPVectori* v = OSQPVectori_malloc(length);
  if (v && length > 0) OSQPVectori_from_raw(v, a);
  return v;
}

OSQPVectorf* OSQPVectorf_copy_new(const OSQPVectorf* a) {
  OSQPVectorf* v = OSQPVectorf_malloc(a->length);
  if (v) OSQPVectorf_copy(v, a);
  return v;
}

void OSQPVectorf_free(OSQPVectorf* a) {
  if (!a) return;
  if (!a->is_view) dev_free(a->d);
  free(a);
}

void OSQPVectori_free(OSQPVectori* a) {
  if (!a) return;
  dev_free(a->d);
  free(a);
}

OSQPVectorf* OSQPVectorf_view(const OSQPVectorf* a, OSQPInt head, OSQPInt length) {
  OSQPVectorf* v = (OSQPVectorf*)calloc(1, sizeof(OSQPVectorf));
  if (!v) return nullptr;
  v->d = a->d ? a->d + head : nullptr;
  v->length = length;
  v->is_view = 1;
  return v;
}

void OSQPVectorf_view_free(OSQPVectorf* a) { free(a); }

OSQPInt OSQPVectorf_length(const OSQPVectorf* a) { return a->length; }
OSQPInt OSQPVectori_length(const OSQPVectori* a) { return a->length; }
OSQPFloat* OSQPVectorf_device_ptr(const OSQPVectorf* a) { return a->d; }

void OSQPVectorf_copy(OSQPVectorf* b, const OSQPVectorf* a) {
  if (!a || !b || a->length <= 0) return;
  HIPCHK(hipMemcpyAsync(b->d, a->d, sizeof(double) * (size_t)a->length, hipMemcpyDeviceToDevice, S()));
}

void OSQPVectorf_from_raw(OSQPVectorf* b, const OSQPFloat* av) {
  if (b->length > 0) upload(b->d, av, sizeof(double) * (size_t)b->length);
}

void OSQPVectori_from_raw(OSQPVectori* b, const OSQPInt* av) {
  if (b->length > 0) upload(b->d, av, sizeof(int) * (size_t)b->length);
}

void OSQPVectorf_to_raw(OSQPFloat* bv, const OSQPVectorf* a) {
  if (a->length > 0) download(bv, a->d, sizeof(double) * (size_t)a->length);
}

void OSQPVectori_to_raw(OSQPInt* bv, const OSQPVectori* a) {
  if (a->length > 0) download(bv, a->d, sizeof(int) * (size_t)a->length);
}

// ---- elementwise math ------------------------------------------------------------
#define EW_LAUNCH(kern, n, ...) \
  do { if ((n) > 0) hipLaunchKernelGGL(kern, dim3(egrid(n)), dim3(kBlock), 0, S(), __VA_ARGS__); } while (0)

void OSQPVectorf_set_scalar(OSQPVectorf* a, OSQPFloat sc) { EW_LAUNCH(k_set, a->length, a->d, sc, a->length); }

void OSQPVectorf_set_scalar_conditional(OSQPVectorf* a, const OSQPVectori* test, OSQPFloat sc_if_neg,
                                        OSQPFloat sc_if_zero, OSQPFloat sc_if_pos) {
  EW_LAUNCH(k_set_cond, a->length, a->d, test->d, sc_if_neg, sc_if_zero, sc_if_pos, a->length);
}

void OSQPVectorf_round_to_zero(OSQPVectorf* a, OSQPFloat tol) { EW_LAUNCH(k_round_to_zero, a->length, a->d, tol, a->length); }

void OSQPVectorf_mult_scalar(OSQPVectorf* a, OSQPFloat sc) { EW_LAUNCH(k_mult_sc, a->length, a->d, sc, a->length); }

void OSQPVectorf_plus(OSQPVectorf* x, const OSQPVectorf* a, const OSQPVectorf* b) {
  EW_LAUNCH(k_plus, a->length, x->d, a->d, b->d, a->length);
}

void OSQPVectorf_minus(OSQPVectorf* x, const OSQPVectorf* a, const OSQPVectorf* b) {
  EW_LAUNCH(k_minus, a->length, x->d, a->d, b->d, a->length);
}

void OSQPVectorf_add_scaled(OSQPVectorf* x, OSQPFloat sca, const OSQPVectorf* a, OSQPFloat scb,
                            const OSQPVectorf* b) {
  EW_LAUNCH(k_add_scaled, x->length, x->d, sca, a->d, scb, b->d, x->length);
}

void OSQPVectorf_add_scaled3(OSQPVectorf* x, OSQPFloat sca, const OSQPVectorf* a, OSQPFloat scb,
                             const OSQPVectorf* b, OSQPFloat scc, const OSQPVectorf* c) {
  int incr = (x == a && sca == 1.0) ? 1 : 0;
  EW_LAUNCH(k_add_scaled3, x->length, x->d, sca, a->d, scb, b->d, scc, c->d, incr, x->length);
}

void OSQPVectorf_ew_prod(OSQPVectorf* c, const OSQPVectorf* a, const OSQPVectorf* b) {
  EW_LAUNCH(k_ew_prod, a->length, c->d, a->d, b->d, a->length);
}

void OSQPVectorf_ew_bound_vec(OSQPVectorf* x, const OSQPVectorf* z, const OSQPVectorf* l, const OSQPVectorf* u) {
  EW_LAUNCH(k_bound, x->length, x->d, z->d, l->d, u->d, x->length);
}

void OSQPVectorf_project_polar_reccone(OSQPVectorf* y, const OSQPVectorf* l, const OSQPVectorf* u, OSQPFloat infval) {
  EW_LAUNCH(k_polar_reccone, y->length, y->d, l->d, u->d, infval, y->length);
}

void OSQPVectorf_ew_reciprocal(OSQPVectorf* b, const OSQPVectorf* a) { EW_LAUNCH(k_reciprocal, a->length, b->d, a->d, a->length); }

void OSQPVectorf_ew_sqrt(OSQPVectorf* a) { EW_LAUNCH(k_sqrt, a->length, a->d, a->length); }

void OSQPVectorf_ew_max_vec(OSQPVectorf* c, const OSQPVectorf* a, const OSQPVectorf* b) {
  EW_LAUNCH(k_max_vec, a->length, c->d, a->d, b->d, a->length);
}

void OSQPVectorf_ew_min_vec(OSQPVectorf* c, const OSQPVectorf* a, const OSQPVectorf* b) {
  EW_LAUNCH(k_min_vec, a->length, c->d, a->d, b->d, a->length);
}

void OSQPVectorf_set_scalar_if_lt(OSQPVectorf* x, const OSQPVectorf* z, OSQPFloat testval, OSQPFloat newval) {
  EW_LAUNCH(k_if_lt, x->length, x->d, z->d, testval, newval, x->length);
}

void OSQPVectorf_set_scalar_if_gt(OSQPVectorf* x, const OSQPVectorf* z, OSQPFloat testval, OSQPFloat newval) {
  EW_LAUNCH(k_if_gt, x->length, x->d, z->d, testval, newval, x->length);
}

// ---- reductions returning host scalars ---------------------------------------------
OSQPFloat OSQPVectorf_norm_inf(const OSQPVectorf* v) {
  RedArgs r{v->d, nullptr, nullptr, nullptr, 0, 0};
  return reduce(R_NORM_INF, r, v->length);
}

OSQPFloat OSQPVectorf_scaled_norm_inf(const OSQPVectorf* Sv, const OSQPVectorf* v) {
  RedArgs r{Sv->d, v->d, nullptr, nullptr, 0, 0};
  return reduce(R_SCALED_NORM_INF, r, v->length);
}

OSQPFloat OSQPVectorf_norm_inf_diff(const OSQPVectorf* a, const OSQPVectorf* b) {
  RedArgs r{a->d, b->d, nullptr, nullptr, 0, 0};
  return reduce(R_NORM_INF_DIFF, r, a->length);
}

OSQPFloat OSQPVectorf_norm_1(const OSQPVectorf* a) {
  RedArgs r{a->d, nullptr, nullptr, nullptr, 0, 0};
  return reduce(R_NORM_1, r, a->length);
}

OSQPFloat OSQPVectorf_dot_prod(const OSQPVectorf* a, const OSQPVectorf* b) {
  RedArgs r{a->d, b->d, nullptr, nullptr, 0, 0};
  return reduce(R_DOT, r, a->length);
}

OSQPFloat OSQPVectorf_dot_prod_signed(const OSQPVectorf* a, const OSQPVectorf* b, OSQPInt sign) {
  RedArgs r{a->d, b->d, nullptr, nullptr, 0, 0};
  if (sign == 1) return reduce(R_DOT_POS, r, a->length);
  if (sign == -1) return reduce(R_DOT_NEG, r, a->length);
  return reduce(R_DOT, r, a->length);
}

void osqp_amd_red_norm_inf(const OSQPVectorf* v, OSQPFloat* dst) {
  RedArgs r{v->d, nullptr, nullptr, nullptr, 0, 0};
  red_queue(R_NORM_INF, r, v->length, dst);
}

void osqp_amd_red_scaled_norm_inf(const OSQPVectorf* Sv, const OSQPVectorf* v, OSQPFloat* dst) {
  RedArgs r{Sv->d, v->d, nullptr, nullptr, 0, 0};
  red_queue(R_SCALED_NORM_INF, r, v->length, dst);
}

void osqp_amd_red_dot_prod(const OSQPVectorf* a, const OSQPVectorf* b, OSQPFloat* dst) {
  RedArgs r{a->d, b->d, nullptr, nullptr, 0, 0};
  red_queue(R_DOT, r, a->length, dst);
}

void osqp_amd_red_dot_prod_signed(const OSQPVectorf* a, const OSQPVectorf* b, OSQPInt sign, OSQPFloat* dst) {
  RedArgs r{a->d, b->d, nullptr, nullptr, 0, 0};
  red_queue(sign == 1 ? R_DOT_POS : (sign == -1 ? R_DOT_NEG : R_DOT), r, a->length, dst);
}

void osqp_amd_red_wait(void) { red_drain(rt()); }

OSQPInt OSQPVectorf_all_leq(const OSQPVectorf* l, const OSQPVectorf* u) {
  RedArgs r{l->d, u->d, nullptr, nullptr, 0, 0};
  return reduce(R_ANY_GT, r, l->length) > 0.0 ? 0 : 1;
}

OSQPInt OSQPVectorf_in_reccone(const OSQPVectorf* y, const OSQPVectorf* l, const OSQPVectorf* u, OSQPFloat infval,
                               OSQPFloat tol) {
  RedArgs r{y->d, l->d, u->d, nullptr, infval, tol};
  return reduce(R_RECCONE_BAD, r, y->length) > 0.0 ? 0 : 1;
}

OSQPInt OSQPVectorf_is_eq(const OSQPVectorf* A, const OSQPVectorf* B, OSQPFloat tol) {
  if (A->length != B->length) return 0;
  RedArgs r{A->d, B->d, nullptr, nullptr, tol, 0};
  return reduce(R_EQ_BAD, r, A->length) > 0.0 ? 0 : 1;
}

OSQPInt OSQPVectorf_ew_bounds_type(OSQPVectori* iseq, const OSQPVectorf* l, const OSQPVectorf* u, OSQPFloat tol,
                                   OSQPFloat infval) {
  int n = iseq->length;
  if (n <= 0) return 0;
  Runtime& R = rt();
  int g = reduce_grid(n);
  const double seq = ++R.red_seq;
  hipLaunchKernelGGL(k_bounds_type, dim3(g), dim3(kBlock), 0, R.stream, iseq->d, l->d, u->d, tol, infval, n,
                     R.d_partials, R.d_ticket, R.dh_scalars, seq);
  if (LAUNCHCHK()) return 0;
  return wait_result(R, seq) > 0.0 ? 1 : 0;
}

void osqp_amd_admm_rhs(OSQPVectorf* xt, OSQPVectorf* zt, OSQPFloat sigma, const OSQPVectorf* x_prev,
                       const OSQPVectorf* q, const OSQPVectorf* rho_inv, const OSQPVectorf* y,
                       const OSQPVectorf* z_prev) {
  const int n = xt->length, m = zt->length;
  EW_LAUNCH(k_admm_rhs, (long long)n + m, n, m, sigma, xt->d, x_prev->d, q->d, zt->d, rho_inv->d, y->d, z_prev->d);
}

void osqp_amd_admm_update(OSQPFloat alpha, OSQPVectorf* x, OSQPVectorf* delta_x, OSQPVectorf* xt,
                          const OSQPVectorf* x_prev, OSQPVectorf* z, OSQPVectorf* y, OSQPVectorf* delta_y,
                          OSQPVectorf* zt, const OSQPVectorf* z_prev, const OSQPVectorf* rho,
                          const OSQPVectorf* rho_inv, const OSQPVectorf* l, const OSQPVectorf* u, OSQPInt next,
                          OSQPFloat sigma, const OSQPVectorf* q, OSQPFloat* w_next) {
  const int n = x->length, m = z->length;
  EW_LAUNCH(k_admm_update, (long long)n + m, n, m, alpha, x->d, delta_x->d, xt->d, x_prev->d, z->d, y->d,
            delta_y->d, zt->d, z_prev->d, rho->d, rho_inv->d, l->d, u->d, next ? 1 : 0, sigma,
            next ? q->d : nullptr, next ? w_next : nullptr);
}

}  // extern "C"

// ---- reductions that allocate nothing ----------------------------------------
// osqp_amd_reduce_noalloc runs up to kMultiMax of the queued reductions' kinds
// in one k_reduce_multi launch, exactly as osqp_amd_red_wait runs them (same
// grids, same order, same bits), but its per-block partials and tickets live in
// a file-scope device array rather than in memory that osqp_algebra_init_libs
// allocates, and its results go to the caller's device memory on the caller's
// stream.  It never calls hipMalloc / hipFree, so code that may not allocate
// (a stream under HIP graph capture, a tight serving loop) can call it.
//
// scratch: kMultiMax blocks of 512 partials, then the tickets (slot 1 + k is
// reduction k's, as in k_reduce_multi).  Each GPU holds its own copy for the
// life of the process; its address comes from hipGetSymbolAddress once per
// device.  Every call on a device shares it, so calls that overlap on one GPU
// (two streams, or two host threads) race.
namespace {
constexpr int kNoAllocPartials = kMultiMax * 512;
constexpr int kNoAllocTickets = 32;  // unsigned slots used: 1 + kMultiMax
static __device__ __align__(256) double scratch[kNoAllocPartials + kNoAllocTickets / 2];

constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mu;

// The tickets must start at zero: every finishing block re-arms its own, but
// a call that never finished (a launch error) would leave one armed.  A fill
// kernel, not hipMemsetAsync, so a captured graph replays it correctly.
__global__ void k_clear_tickets(unsigned* t, int n) {
  if ((int)threadIdx.x < n) t[threadIdx.x] = 0u;
}

double* scratch_ptr_for_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    double* p = nullptr;
    if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" OSQPInt osqp_amd_reduce_noalloc(const OSQPAmdReduction* items, OSQPInt count, OSQPFloat* out,
                                           void* stream) {
  if (count < 1 || count > kMultiMax) {
    fprintf(stderr, "osqp_amd_reduce_noalloc: %d reductions requested; the scratch holds 1 to %d per call\n",
            (int)count, kMultiMax);
    return OSQP_DATA_VALIDATION_ERROR;
  }
  if (!items || !out) {
    fprintf(stderr, "osqp_amd_reduce_noalloc: null items or out\n");
    return OSQP_DATA_VALIDATION_ERROR;
  }
  RedBatch B;
  int gmax = 1;
  for (int k = 0; k < count; ++k) {
    const OSQPAmdReduction& it = items[k];
    int op = -1;
    switch (it.kind) {
      case OSQP_AMD_RED_NORM_INF: op = R_NORM_INF; break;
      case OSQP_AMD_RED_SCALED_NORM_INF: op = R_SCALED_NORM_INF; break;
      case OSQP_AMD_RED_NORM_INF_DIFF: op = R_NORM_INF_DIFF; break;
      case OSQP_AMD_RED_NORM_1: op = R_NORM_1; break;
      case OSQP_AMD_RED_DOT: op = R_DOT; break;
      case OSQP_AMD_RED_DOT_POS: op = R_DOT_POS; break;
      case OSQP_AMD_RED_DOT_NEG: op = R_DOT_NEG; break;
      default: break;
    }
    const bool two = op == R_SCALED_NORM_INF || op == R_NORM_INF_DIFF || op == R_DOT || op == R_DOT_POS ||
                     op == R_DOT_NEG;
    if (op < 0 || it.n < 0 || (it.n > 0 && (!it.a || (two && !it.b)))) {
      fprintf(stderr, "osqp_amd_reduce_noalloc: item %d: bad kind %d, length %d or null operand\n", k, (int)it.kind,
              (int)it.n);
      return OSQP_DATA_VALIDATION_ERROR;
    }
    RedTask& T = B.t[k];
    T.op = op;
    T.n = (int)it.n;
    T.g = red_grid((int)it.n);  // 1 block for n = 0: it writes 0, as the host-side path returns
    T.pad = 0;
    T.r = RedArgs{it.a, two ? it.b : nullptr, nullptr, nullptr, 0, 0};
    gmax = T.g > gmax ? T.g : gmax;
  }
  for (int k = count; k < kMultiMax; ++k) B.t[k] = RedTask{};
  double* ptr = scratch_ptr_for_current_device();
  if (!ptr) {
    fprintf(stderr, "osqp_amd_reduce_noalloc: no device address for the scratch array on this device\n");
    return OSQP_ALGEBRA_LOAD_ERROR;
  }
  unsigned* tickets = reinterpret_cast<unsigned*>(ptr + kNoAllocPartials);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_clear_tickets, dim3(1), dim3(64), 0, s, tickets, 1 + kMultiMax);
  hipLaunchKernelGGL(k_reduce_multi<true>, dim3(gmax, count), dim3(kBlock), 0, s, B, ptr, tickets, out, 0.0);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "osqp_amd_reduce_noalloc: launch failed: %s\n", hipGetErrorString(e));
    return OSQP_ALGEBRA_LOAD_ERROR;
  }
  return 0;
}
