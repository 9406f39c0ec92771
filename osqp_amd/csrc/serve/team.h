// team.h — one OSQP problem solved by one workgroup (a "team").
//
// The batched engine keeps many small problems resident in HBM and runs one
// team per problem.  Everything here is device code that restates, for one
// problem held by one workgroup, the host core's ADMM step (src/osqp_api.c
// osqp_solve loop :698-903, src/auxil.c, src/scaling.c) and the PCG linear
// solver of the single-problem engine (tolerance schedule
// algebra/cuda/lin_sys/indirect/cuda_pcg_interface.cu:32-64, CG
// cuda_pcg.cu:113-208, with FP64 defect correction around an FP32 or FP64 CG).
//
// Results do not depend on the team size (threads per workgroup) or on the
// other problems in flight:
//   * a matrix's rows of at most 64 entries are packed, in row order, into
//     fixed groups of 64 lanes x 8 entries (no row crosses a group); a group is
//     reduced by one wavefront in a fixed order (each lane adds its 8 entries
//     in sequence, a fixed-shape scan joins the rows that cross lanes),
//     whatever wavefront takes it; longer rows are reduced by one wavefront
//     each, also in a fixed order;
//   * a sum / max over a vector is formed over fixed 256-element blocks (each
//     block by one wavefront: 4 elements per lane in index order, then a fixed
//     butterfly) and the block results are added in block order.
// So the engine's tuner can change team sizes freely, and a problem's bits are
// the same alone and in any mix.
#pragma once

#include <math.h>
#include <stdint.h>

#include "../hip/hip_internal.h"

namespace osqp_serve {

using osqp_hip::kWave;

constexpr int kRedBlock = 256;     // elements per reduction block
constexpr int kMaxRedBlocks = 512; // vectors up to 131072 entries
constexpr int kLongRow = 64;       // rows longer than this go to whole wavefronts
constexpr int kGroupPer = 8;       // entries per lane of a row group
constexpr int kGroupSlots = kWave * kGroupPer;  // entries per row group
constexpr int kLongPer = 16;     // strides of a long row whose loads are issued together
constexpr int kLongAlign = 8;    // compact long rows start at multiples of 8 entries
constexpr int kLongVec = 4;      // compact long-row entries per lane per load (float4 + 4 x 16 bits)
constexpr int kLongVecRounds = 16;  // compact loads per lane issued together (sliced kernel)
constexpr int kLongVecRoundsLean = 8;  // the same in the kernel without slices
constexpr int kEllMin = 8;         // FP32 CG: rows of kEllMin..kLongRow entries go to slices
constexpr int kEllMinRows = 64;    // ... when a matrix has at least this many of them
constexpr int kEllVec = 4;         // slice entries per lane per load (float4 + 4 x 16 bits)
constexpr int kEllRounds = 8;      // slice loads per lane issued together
constexpr int kMaxTeamWaves = 8;   // teams of up to 512 threads
constexpr int kTeamLdsBytes = 32768;  // least dynamic LDS of a team (more where fewer teams share a CU)

// phases of a resident problem
enum : int { kPhaseIdle = 0, kPhaseSetup = 1, kPhaseUpdate = 2, kPhaseRun = 3, kPhaseDone = 4 };

// status values (include/public/osqp_api_constants.h)
enum : int {
  kSolved = 1,
  kSolvedInaccurate = 2,
  kPrimalInfeasible = 3,
  kPrimalInfeasibleInaccurate = 4,
  kDualInfeasible = 5,
  kDualInfeasibleInaccurate = 6,
  kMaxIterReached = 7,
  kTimeLimit = 8,
  kNonCvx = 9,
  kUnsolved = 11
};

constexpr double kInfty = 1e30;
constexpr double kMinScaling = 1e-4;
constexpr double kMaxScaling = 1e4;
constexpr double kDivisionTol = 1.0 / kInfty;
constexpr double kRhoMin = 1e-6;
constexpr double kRhoMax = 1e6;
constexpr double kRhoTol = 1e-4;
constexpr double kRhoEqOverIneq = 1e3;
constexpr double kZeroDeadzone = 1e-15;
constexpr double kCgTolMin = 1e-7;
constexpr int kCheckTermDefault = 25;  // OSQP_CHECK_TERMINATION
// OSQP_NAN = (OSQPFloat)0x7fc00000 (include/public/osqp_api_constants.h:193):
// the number 2143289344.0, not an IEEE NaN
constexpr double kOsqpNan = 2143289344.0;
// adaptive_rho modes (include/public/osqp_api_constants.h:125-128)
enum : int { kAdaptOff = 0, kAdaptIters = 1, kAdaptTime = 2, kAdaptKkt = 3 };

// CSR matrix with its row groups (rows of 1..kLongRow entries), long-row and
// empty-row lists.  Slot (g, lane, e) of the groups sits at
// g * kGroupSlots + lane * kGroupPer + e and holds the group's entry
// lane * kGroupPer + e (row order, then column order), or padding
// (grow = -1) after the group's last entry.  A lane's slots are
// contiguous, so it loads them with 16-byte vector loads.  The host writes
// the rows and the entries' offsets within them; fill_group_values derives
// the columns and values from the CSR arrays.
struct BCsr {
  int rows;
  int ng;       // row groups
  int nlong;    // rows longer than kLongRow
  int nempty;   // rows without entries
  int* rp;
  int* ci;
  double* val;
  int* longs;
  int* empty;
  unsigned char* goff;  // ng * kGroupSlots: the slot's entry's offset within its row
  int* grow;    // its row, -1 padding
  int* gci;     // its column (0 for padding)
  double* gval; // its value, copied from val after scaling (0 for padding)
  // compact copy of the long rows for the FP32 CG (null when absent: FP64
  // engine, or 65536 or more columns): 6 bytes an entry instead of 12; the
  // values are (float) of val, so the products are the same
  int* loff;             // nlong + 1: start of each long row in the compact long-row arrays
  unsigned short* lci16; // long-row entries (rows start at multiples of kLongAlign, zero padding)
  float* lv32;           // their (float) values, 0 for padding
  // slices of the medium rows for the FP32 CG (null when absent; same
  // conditions, and a matrix with fewer than kEllMinRows medium rows has
  // none): the FP32 CG walks groups [0, ng32), which then hold only the
  // short rows, and takes the medium rows from the slices, one row per lane
  // summed in CSR order.  Slice s: (row, length) of lane l's row at
  // erl[2 (kWave s + l)] (row -1 for padding), entries [eoff[s], eoff[s + 1]);
  // entry k of the lane's row at
  // eoff[s] + (k / kEllVec) kWave kEllVec + lane kEllVec + k % kEllVec.
  int ng32;              // groups the FP32 CG walks (ng without slices)
  int nslice;
  int* eoff;             // nslice + 1
  int* erl;              // nslice * kWave (row, length) pairs
  unsigned short* eci16; // columns (0 for padding)
  float* ev32;           // (float) values, 0 for padding
};

// scalar state that survives between epochs (stored with the problem)
struct BState {
  int phase;
  int iter;          // ADMM iterations of the current solve
  int swapped;       // parity of the x/x_prev and z/z_prev buffer swaps
  int status;
  int rho_updates;
  int can_check;     // the last iteration ran a termination check
  int ax_valid;      // ax64 == A x64
  int cg_iter;       // CG iterations of the last linear solve
  int cg_zero_iters; // consecutive linear solves with no CG iteration
  int cg_solves;
  long long cg_total;
  double cg_eps_prev, cg_red_factor;
  double c, cinv;    // cost scaling
  double rho;        // current scalar rho
  double obj_val, dual_obj_val, prim_res, dual_res, gap, primdual_int, rel_kkt, rho_estimate;
  double scaled_prim_res, scaled_dual_res, scaled_gap, xtPx, qtx, SC, last_rel_kkt;
  // device wall-clock ticks: start of the solve's run time (setup start for a
  // new problem, as the reference counts setup_time + solve time), start of
  // the ADMM loop, and the setup phase's duration
  long long t_start, t_solve, setup_ticks;
  int ad_interval;   // rho adaptation interval (fixed by the time mode once)
};

// settings copied into every problem
struct BSettings {
  double rho0, sigma, alpha, eps_abs, eps_rel, eps_pinf, eps_dinf;
  double ad_tol, cg_tol_fraction, ad_fraction;
  long long tl_ticks;  // time_limit in device wall-clock ticks, 0 = none
  int max_iter, scaling, check_term, adaptive, ad_interval, scaled_term, check_gap, warm;
  int cg_max_iter, cg_red_thresh;
};

// result record written to pinned host memory when a solve ends
struct BResult {
  int status, iter, rho_updates, done;
  double obj_val, dual_obj_val, prim_res, dual_res, gap, rho_estimate, primdual_int, rel_kkt;
  long long cg_iters;
};

// new matrix values of a resident problem (osqp_update_data_mat), in host-
// mapped memory: np values for slots of P's full CSR, then na values for A's
// CSR slots with, at the same index, the A' slot of the same entry
struct BMatUpd {
  int np, na;
  const int* pslot;
  const double* pval;
  const int* aslot;
  const int* tslot;
  const double* aval;
};

// one resident problem: device pointers into its arena block
struct BProb {
  int n, m;
  BSettings set;
  BState st;
  BCsr P, A, At;   // P full (both triangles), A, A' (= A's CSC)
  int* Pdiag;      // position of P's diagonal in each row, -1 if none
  int* ctype;      // constraint type: -1 loose, 0 inequality, 1 equality
  double *q, *l, *u, *D, *Dinv, *E, *Einv, *rho, *rhoinv;
  double *xa, *xb, *za, *zb, *y, *xt, *zt, *dx, *dy, *Ax, *Px, *Aty, *tn, *tm, *tn2;
  double *x64, *r64, *rhs64, *w64, *ax64, *prec;
  void *cd, *cp, *cKp, *cr, *cy, *cw, *cdinv;  // CG vectors (FP32 or FP64)
  // host-mapped staging (read in the setup / update phases) and results
  const char* stage;     // static data image, same layout as the arena prefix
  long long stage_bytes;
  char* arena_static;    // destination of the static image
  const double* upd_q;   // new q for an update request (raw, unscaled) or null
  const double* warm_x;  // warm start (raw) or null
  const double* warm_y;
  double* out_x;         // results (pinned host)
  double* out_y;
  double* out_pcert;     // primal / dual infeasibility certificates
  double* out_dcert;
  BResult* out;
};

// pointers with their address space stated (global: the arena and host-
// mapped staging; LDS: the workgroup's dynamic LDS), so that loads are
// global_load / ds_read rather than flat (a flat load's wait also waits for
// every LDS access issued before it, and the other way round)
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
using lptr = __attribute__((address_space(3))) T*;

// ---- DPP lane moves (GFX9 data-parallel primitives) --------------------------
// v_mov_b32_dpp: each lane reads a register of another lane of its wavefront
// in the same instruction, without an LDS round trip.  Lanes without a source
// (or outside row_mask) keep `old`.  Controls: row_shr:d within rows of 16
// lanes, wave_shr:1 / wave_shl:1 across the wavefront, row_bcast:15 (lane 15
// of each row to the next row), row_bcast:31 (lane 31 to rows 2 and 3).
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138, kDppBcast15 = 0x142, kDppBcast31 = 0x143;

template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp_move(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ float dpp_move(float old, float v) {
  return __int_as_float(dpp_move<CTRL, ROWS>(__float_as_int(old), __float_as_int(v)));
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_move(double old, double v) {
  const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
  const int lo = dpp_move<CTRL, ROWS>((int)(o & 0xffffffff), (int)(x & 0xffffffff));
  const int hi = dpp_move<CTRL, ROWS>((int)(o >> 32), (int)(x >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// one step of the inclusive scan of the maps x -> b + a x (a in {0, 1}): the
// lane's map after the one read from another lane (identity: a = 1, b = 0)
template <int CTRL, int ROWS = 0xf, typename TX>
__device__ __forceinline__ void scan_step(int& a, TX& b) {
  const int ap = dpp_move<CTRL, ROWS>(1, a);
  const TX bp = dpp_move<CTRL, ROWS>((TX)0, b);
  if (a) b = b + bp;
  a = a & ap;
}

// ---- team primitives -------------------------------------------------------
struct Team {
  int tid, T, lane, wv, W;
  char* lds;       // dynamic LDS of the workgroup (small problems' CG vectors)
  int lds_bytes;
  double* red;   // LDS, kMaxRedBlocks doubles
  double* red2;  // LDS, kMaxRedBlocks doubles (a second reduction in the same pass)
  int* flag;     // LDS word for decisions that thread 0 takes for the team
  __device__ __forceinline__ void sync() const { __syncthreads(); }
  // the team's common value of a decision taken by thread 0 (reads of the
  // device clock differ between threads)
  __device__ __forceinline__ int decide(int v0) const {
    if (tid == 0) *flag = v0;
    __syncthreads();
    const int v = *flag;
    __syncthreads();
    return v;
  }
};

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

template <typename V>
struct KahanAcc {
  V s = 0, c = 0;
  __device__ __forceinline__ void add(V v) {
    const V y = v - c;
    const V t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};
template <>
struct KahanAcc<double> {
  double s = 0;
  __device__ __forceinline__ void add(double v) { s += v; }
  __device__ __forceinline__ double value() const { return s; }
};

// sum (MAX = false) or max (MAX = true, of non-negative terms) of f(i), i < n;
// block-structured, independent of the team size.  Ends with every thread
// holding the result; contains two barriers.
template <bool MAX, typename F>
__device__ __forceinline__ double team_reduce(const Team& t, int n, F f) {
  const int nb = (n + kRedBlock - 1) / kRedBlock;
  for (int b = t.wv; b < nb; b += t.W) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < kRedBlock / kWave; ++j) {
      const int i = b * kRedBlock + j * kWave + t.lane;
      if (i < n) {
        const double v = f(i);
        acc = MAX ? dmax(acc, v) : acc + v;
      }
    }
    acc = MAX ? osqp_hip::wave_max(acc) : osqp_hip::wave_sum(acc);
    if (t.lane == 0) t.red[b] = acc;
  }
  t.sync();
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s = MAX ? dmax(s, t.red[b]) : s + t.red[b];
  t.sync();
  return s;
}

// dot product in the CG precision V (Kahan-compensated in FP32), same blocking
template <typename V, typename F>
__device__ __forceinline__ double team_dot(const Team& t, int n, F f) {
  const int nb = (n + kRedBlock - 1) / kRedBlock;
  for (int b = t.wv; b < nb; b += t.W) {
    KahanAcc<V> acc;
#pragma unroll
    for (int j = 0; j < kRedBlock / kWave; ++j) {
      const int i = b * kRedBlock + j * kWave + t.lane;
      if (i < n) acc.add(f(i));
    }
    const double v = osqp_hip::wave_sum(acc.value());
    if (t.lane == 0) t.red[b] = v;
  }
  t.sync();
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += t.red[b];
  t.sync();
  return s;
}

// One pass over the same fixed 256-element blocks as team_dot and
// team_reduce: body(i) runs once for every element (it may update it) and
// returns the pair (dot term, max term); the result is (team_dot of the
// first, team_reduce<true> of the second) with the bits of the two separate
// reductions.  Four barriers fewer than body loop + team_dot + team_reduce.
template <typename V, typename F>
__device__ __forceinline__ void team_dot_max(const Team& t, int n, F body, double* dot, double* mx) {
  const int nb = (n + kRedBlock - 1) / kRedBlock;
  for (int b = t.wv; b < nb; b += t.W) {
    KahanAcc<V> acc;
    double m = 0.0;
#pragma unroll
    for (int j = 0; j < kRedBlock / kWave; ++j) {
      const int i = b * kRedBlock + j * kWave + t.lane;
      if (i < n) {
        V a;
        double c;
        body(i, a, c);
        acc.add(a);
        m = dmax(m, c);
      }
    }
    const double v = osqp_hip::wave_sum(acc.value());
    m = osqp_hip::wave_max(m);
    if (t.lane == 0) {
      t.red[b] = v;
      t.red2[b] = m;
    }
  }
  t.sync();
  double s = 0.0, x = 0.0;
  for (int b = 0; b < nb; ++b) {
    s += t.red[b];
    x = dmax(x, t.red2[b]);
  }
  t.sync();
  *dot = s;
  *mx = x;
}

// y-rows of a CSR product: for each row r, epi(r, sum_k elem(val_k, col_k)),
// epi called exactly once per row by one lane.  Row groups go to wavefronts
// (one group per wavefront at a time, the next group's slots loaded while
// this one is reduced):
//   1. each lane loads its 8 slots, forms the products and adds them in order,
//      row by row: a row wholly inside the lane is finished there; the lane's
//      first row (head) and last row (tail) may continue across lanes;
//   2. C_l, the part of lane l's head row from lanes before l, follows
//      C_l = cont_l ? ts_{l-1} + (single_{l-1} ? C_{l-1} : 0) : 0 (ts: tail sum,
//      single: the lane holds one row), an affine recurrence solved by a
//      6-step DPP scan of fixed shape (rows of 16 lanes, then row broadcasts);
//   3. the lane where a row ends adds C_l to its part and runs the epilogue.
// Long rows take one wavefront each; empty rows get epi(r, 0).  The matrix is
// taken by value so its pointers live in registers.  Caller syncs.
//
// With a compact copy of the long rows (FP32 CG, fewer than 65536 columns) a
// long row's lane adds 4 consecutive entries per load ((float) values, 16-bit
// columns) in order, chunk after chunk.

// steps 1-3 for one group, from each slot's row (-1: padding; padding slots
// are a lane's last ones) and product.  Branch-free up to the epilogues: the
// running sum of each slot's row within the lane (the same additions in the
// same order as a row-by-row loop), the lane's tail, the DPP scan, then one
// epilogue per row end.
template <typename TX, typename Epi>
__device__ __forceinline__ void group_rows(const int lane, Epi& epi, const int (&rr)[kGroupPer],
                                           const TX (&v)[kGroupPer]) {
  TX seg[kGroupPer];
  seg[0] = v[0];
#pragma unroll
  for (int e = 1; e < kGroupPer; ++e) seg[e] = rr[e] == rr[e - 1] ? seg[e - 1] + v[e] : v[e];
  const int hr = rr[0];
  int tr = hr;
  TX ts = seg[0];
#pragma unroll
  for (int e = 1; e < kGroupPer; ++e) {
    tr = rr[e] >= 0 ? rr[e] : tr;
    ts = rr[e] >= 0 ? seg[e] : ts;
  }
  const bool multi = hr >= 0 && tr != hr;
  // lane l's head continues lane l-1's tail
  const int prev_tr = dpp_move<kDppWaveShr1>(-2, tr);
  const TX prev_ts = dpp_move<kDppWaveShr1>((TX)0, ts);
  const int prev_multi = dpp_move<kDppWaveShr1>(0, (int)multi);
  const bool cont = lane > 0 && hr >= 0 && hr == prev_tr;
  int a = (cont && !prev_multi) ? 1 : 0;
  TX b = cont ? prev_ts : (TX)0;
  // inclusive scan: within rows of 16 lanes, then the end of rows 0 / 2
  // into rows 1 / 3, then the end of row 1 into rows 2 and 3
  scan_step<kDppRowShr1>(a, b);
  scan_step<kDppRowShr2>(a, b);
  scan_step<kDppRowShr4>(a, b);
  scan_step<kDppRowShr8>(a, b);
  scan_step<kDppBcast15, 0xa>(a, b);
  scan_step<kDppBcast31, 0xc>(a, b);
  const int next_cont = dpp_move<kDppWaveShl1>(0, (int)cont);
  const bool ends_here = lane == kWave - 1 || !next_cont;
  // a row ends at slot e when the next slot holds another row (or padding);
  // the lane's tail row only when the next lane does not continue it.  The
  // head row's value carries b (zero unless it continues from earlier lanes).
#pragma unroll
  for (int e = 0; e < kGroupPer; ++e) {
    const int r = rr[e];
    const bool last = e == kGroupPer - 1 || rr[e + 1 < kGroupPer ? e + 1 : e] != r;
    const bool is_tail = r == tr;
    if (r >= 0 && last && (!is_tail || ends_here)) epi(r, r == hr ? b + seg[e] : seg[e]);
  }
}

// SL: the kernel with slices (the large problems' buckets).  Without them the
// slice path is not compiled at all: its registers, present even when unused,
// cost the small problems' group loops 40 % more time.
template <typename TX, bool SL = false, typename Elem, typename Epi>
__device__ __forceinline__ void team_rows(const Team& t, const BCsr M, Elem elem, Epi epi) {
  const int lane = t.lane;
  static_assert(kGroupPer == 8, "a lane's slots are loaded as two int4 / four double2 vectors");
  typedef int i32x4 __attribute__((ext_vector_type(4)));
  typedef double f64x2 __attribute__((ext_vector_type(2)));
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  bool compact = false, sliced = false;
  if constexpr (sizeof(TX) == 4) {
    compact = M.loff != nullptr;
    if constexpr (SL) sliced = M.eoff != nullptr;
  }
  const int ng = sliced ? M.ng32 : M.ng;
  {
    const gptr<const int> grow = (gptr<const int>)M.grow;
    const gptr<const int> gci = (gptr<const int>)M.gci;
    const gptr<const double> gval = (gptr<const double>)M.gval;
    struct Slots {
      i32x4 r[2], c[2];
      f64x2 v[4];
    };
    auto load = [&](int g, Slots& L) {
      const int base = g * kGroupSlots + lane * kGroupPer;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        L.r[h] = *(gptr<const i32x4>)(grow + base + 4 * h);
        L.c[h] = *(gptr<const i32x4>)(gci + base + 4 * h);
      }
#pragma unroll
      for (int h = 0; h < 4; ++h) L.v[h] = *(gptr<const f64x2>)(gval + base + 2 * h);
    };
    auto reduce = [&](const Slots& L) {
      int rr[kGroupPer], cc[kGroupPer];
      double vv[kGroupPer];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        rr[4 * h + 0] = L.r[h].x;
        rr[4 * h + 1] = L.r[h].y;
        rr[4 * h + 2] = L.r[h].z;
        rr[4 * h + 3] = L.r[h].w;
        cc[4 * h + 0] = L.c[h].x;
        cc[4 * h + 1] = L.c[h].y;
        cc[4 * h + 2] = L.c[h].z;
        cc[4 * h + 3] = L.c[h].w;
      }
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        vv[2 * h + 0] = L.v[h].x;
        vv[2 * h + 1] = L.v[h].y;
      }
      // every slot's product (padding: column 0, in bounds), then the select,
      // so the gathers are issued together
      TX pr[kGroupPer], v[kGroupPer];
#pragma unroll
      for (int e = 0; e < kGroupPer; ++e) pr[e] = (TX)elem(vv[e], cc[e]);
#pragma unroll
      for (int e = 0; e < kGroupPer; ++e) v[e] = rr[e] >= 0 ? pr[e] : (TX)0;
      group_rows<TX>(lane, epi, rr, v);
    };
    if (t.wv < ng) {
      Slots cur;
      load(t.wv, cur);
      for (int g = t.wv;;) {
        const int gn = g + t.W;
        Slots nxt;
        if (gn < ng) load(gn, nxt);
        reduce(cur);
        if (gn >= ng) break;
        cur = nxt;
        g = gn;
      }
    }
  }
  const gptr<const int> empty = (gptr<const int>)M.empty;
  for (int j = t.tid; j < M.nempty; j += t.T) epi(empty[j], (TX)0);
  const gptr<const int> rp = (gptr<const int>)M.rp;
  if constexpr (SL) if (sliced) {
    // slices: lane l sums its row's entries in order, kEllVec per load.  A
    // wave takes slices wv, wv + W, ...; lane i holds the offsets of the
    // i-th of them (windows of kWave slices), and the next slice's first
    // kEllRounds loads of each kind are in flight while this one is added
    // (loads past the slice's width read its first chunk, not added; entries
    // past the row's length are padding, not added).  Wider slices load the
    // rest chunk by chunk.
    const gptr<const int> eoff = (gptr<const int>)M.eoff;
    const gptr<const int> erl = (gptr<const int>)M.erl;
    const gptr<const float> ev32 = (gptr<const float>)M.ev32;
    const gptr<const unsigned short> eci16 = (gptr<const unsigned short>)M.eci16;
    typedef int i32x2 __attribute__((ext_vector_type(2)));
    constexpr int kChunk = kWave * kEllVec;
    struct Slice {
      f32x4 v[kEllRounds];
      u32x2 c[kEllRounds];
      i32x2 rl;
      int b0, nch;
    };
    auto sload = [&](int sl, int b0, int b1, Slice& L) {
      L.b0 = b0;
      L.nch = (b1 - b0) / kChunk;
      L.rl = *(gptr<const i32x2>)(erl + 2 * (sl * kWave + lane));
#pragma unroll
      for (int i = 0; i < kEllRounds; ++i) {
        const int k = b0 + (i < L.nch ? i : 0) * kChunk + lane * kEllVec;
        L.v[i] = *(gptr<const f32x4>)(ev32 + k);
        L.c[i] = *(gptr<const u32x2>)(eci16 + k);
      }
    };
    auto add4 = [&](TX& acc, const f32x4& v, const u32x2& c, int e0, int len) {
#pragma unroll
      for (int q = 0; q < kEllVec; ++q) {
        const unsigned w = c[q >> 1];
        const int col = (int)((q & 1) ? (w >> 16) : (w & 0xFFFFu));
        const TX pr = (TX)elem((double)v[q], col);
        if (e0 + q < len) acc += pr;
      }
    };
    auto sreduce = [&](const Slice& L) {
      TX acc = 0;
      const int len = L.rl.y;
#pragma unroll
      for (int i = 0; i < kEllRounds; ++i) add4(acc, L.v[i], L.c[i], i * kEllVec, len);
      for (int j = kEllRounds; j < L.nch; ++j) {
        const int k = L.b0 + j * kChunk + lane * kEllVec;
        const f32x4 v = *(gptr<const f32x4>)(ev32 + k);
        const u32x2 c = *(gptr<const u32x2>)(eci16 + k);
        add4(acc, v, c, j * kEllVec, len);
      }
      if (L.rl.x >= 0) epi(L.rl.x, acc);
    };
    for (int base = t.wv; base < M.nslice; base += kWave * t.W) {
      const int mine = base + lane * t.W;
      const int o0 = mine < M.nslice ? eoff[mine] : 0;
      const int o1 = mine < M.nslice ? eoff[mine + 1] : 0;
      const int cnt = (M.nslice - base + t.W - 1) / t.W < kWave ? (M.nslice - base + t.W - 1) / t.W : kWave;
      Slice cur;
      sload(base, __builtin_amdgcn_readfirstlane(o0), __builtin_amdgcn_readfirstlane(o1), cur);
      for (int i = 0;;) {
        const int in = i + 1;
        Slice nxt;
        if (in < cnt) sload(base + in * t.W, __builtin_amdgcn_readlane(o0, in), __builtin_amdgcn_readlane(o1, in), nxt);
        sreduce(cur);
        if (in >= cnt) break;
        cur = nxt;
        i = in;
      }
    }
  }
  const gptr<const int> longs = (gptr<const int>)M.longs;
  if (compact) {
    // compact long rows: lane l adds entries s + 4 (64 i + l) + q, q = 0..3,
    // for i = 0, 1, ... in order (kLongVecRounds loads of each kind in
    // flight; chunks that start past the row's end are not loaded, and a
    // lane past the end of the last chunk reads the row's first entries, not
    // added)
    const gptr<const int> loff = (gptr<const int>)M.loff;
    const gptr<const float> lv32 = (gptr<const float>)M.lv32;
    const gptr<const unsigned short> lci16 = (gptr<const unsigned short>)M.lci16;
    constexpr int kChunk = kWave * kLongVec;
    for (int j = t.wv; j < M.nlong; j += t.W) {
      const int r = longs[j];
      const int s = loff[j], e = s + (rp[r + 1] - rp[r]);
      TX acc = 0;
      constexpr int kRounds = SL ? kLongVecRounds : kLongVecRoundsLean;
      for (int k0 = s; k0 < e; k0 += kRounds * kChunk) {
        f32x4 vv[kRounds];
        u32x2 cw[kRounds];
#pragma unroll
        for (int i = 0; i < kRounds; ++i) {
          const int k = k0 + i * kChunk + lane * kLongVec;
          const int kk = k < e ? k : s;
          vv[i] = f32x4{0.f, 0.f, 0.f, 0.f};
          cw[i] = u32x2{0u, 0u};
          if (k0 + i * kChunk < e) {
            vv[i] = *(gptr<const f32x4>)(lv32 + kk);
            cw[i] = *(gptr<const u32x2>)(lci16 + kk);
          }
        }
#pragma unroll
        for (int i = 0; i < kRounds; ++i) {
          const int k = k0 + i * kChunk + lane * kLongVec;
#pragma unroll
          for (int q = 0; q < kLongVec; ++q) {
            const unsigned w = cw[i][q >> 1];
            const int c = (int)((q & 1) ? (w >> 16) : (w & 0xFFFFu));
            const TX pr = (TX)elem((double)vv[i][q], c);
            if (k + q < e) acc += pr;
          }
        }
      }
      acc = osqp_hip::wave_sum(acc);
      if (t.lane == 0) epi(r, acc);
    }
    return;
  }
  // long rows: lane l adds entries s + l, s + l + 64, ... in order; the loads
  // of kLongPer consecutive strides are issued together, the last round's
  // past the row's end at the row's first entry (in bounds, not added)
  const gptr<const double> val = (gptr<const double>)M.val;
  const gptr<const int> ci = (gptr<const int>)M.ci;
  for (int j = t.wv; j < M.nlong; j += t.W) {
    const int r = longs[j];
    const int s = rp[r], e = rp[r + 1];
    TX acc = 0;
    for (int k0 = s + lane; k0 < e; k0 += kLongPer * kWave) {
      double vv[kLongPer];
      int cc[kLongPer];
#pragma unroll
      for (int i = 0; i < kLongPer; ++i) {
        const int k = k0 + i * kWave < e ? k0 + i * kWave : s;
        vv[i] = val[k];
        cc[i] = ci[k];
      }
      TX pr[kLongPer];
#pragma unroll
      for (int i = 0; i < kLongPer; ++i) pr[i] = (TX)elem(vv[i], cc[i]);
#pragma unroll
      for (int i = 0; i < kLongPer; ++i)
        if (k0 + i * kWave < e) acc += pr[i];
    }
    acc = osqp_hip::wave_sum(acc);
    if (t.lane == 0) epi(r, acc);
  }
}

// y = M x (FP64); caller syncs before reading y
__device__ __forceinline__ void team_spmv(const Team& t, const BCsr M, const double* __restrict__ x,
                                          double* __restrict__ y) {
  team_rows<double>(
      t, M, [=](double v, int c) { return v * x[c]; }, [=](int r, double s) { y[r] = s; });
}

// ---- the ADMM step of one problem -------------------------------------------
// The problem descriptor is shared by the team (LDS, read-only here); the
// scalar state S is a private copy in every thread, updated identically by all
// of them from team-wide reductions (no shared read-modify-write).
template <typename V, bool SL = false>
struct Solver {
  const Team& t;
  const BProb& p;
  BState S;
  __device__ __forceinline__ Solver(const Team& tt, const BProb& pp) : t(tt), p(pp), S(pp.st) {}

  __device__ __forceinline__ double* x() { return S.swapped ? p.xb : p.xa; }
  __device__ __forceinline__ double* xprev() { return S.swapped ? p.xa : p.xb; }
  __device__ __forceinline__ double* z() { return S.swapped ? p.zb : p.za; }
  __device__ __forceinline__ double* zprev() { return S.swapped ? p.za : p.zb; }
  __device__ __forceinline__ bool unscaled_term() const { return p.set.scaling && !p.set.scaled_term; }

  __device__ __forceinline__ double norm_inf(const double* v, int n) {
    return team_reduce<true>(t, n, [&](int i) { return fabs(v[i]); });
  }
  __device__ __forceinline__ double scaled_norm_inf(const double* s, const double* v, int n) {
    return team_reduce<true>(t, n, [&](int i) { return fabs(s[i] * v[i]); });
  }
  __device__ __forceinline__ double dot(const double* a, const double* b, int n) {
    return team_reduce<false>(t, n, [&](int i) { return a[i] * b[i]; });
  }

  // ---- scaling (src/scaling.c:48-164) ----------------------------------------
  __device__ __forceinline__ static double limit_scalar(double v) {
    v = v < kMinScaling ? 1.0 : v;
    return v > kMaxScaling ? kMaxScaling : v;
  }
  __device__ __forceinline__ static double limit_elem(double v) {
    v = v < kMinScaling ? 1.0 : v;
    return v > kMaxScaling ? kMaxScaling : v;
  }
  // P entry (row r, col c) scaled as the triangular storage does:
  // (v * s[min]) * s[max]
  __device__ __forceinline__ void scale_P(const double* s) {
    for (int r = t.tid; r < p.n; r += t.T)
      for (int k = p.P.rp[r]; k < p.P.rp[r + 1]; ++k) {
        const int c = p.P.ci[k];
        const int lo = c < r ? c : r, hi = c < r ? r : c;
        p.P.val[k] = (p.P.val[k] * s[lo]) * s[hi];
      }
  }
  // A (i, j) and A' (j, i): (v * E_i) * D_j in both copies
  __device__ __forceinline__ void scale_A(const double* e, const double* d) {
    for (int r = t.tid; r < p.m; r += t.T)
      for (int k = p.A.rp[r]; k < p.A.rp[r + 1]; ++k) p.A.val[k] = (p.A.val[k] * e[r]) * d[p.A.ci[k]];
    for (int r = t.tid; r < p.n; r += t.T)
      for (int k = p.At.rp[r]; k < p.At.rp[r + 1]; ++k) p.At.val[k] = (p.At.val[k] * e[p.At.ci[k]]) * d[r];
  }
  __device__ __forceinline__ void scale_data() {
    const int n = p.n, m = p.m;
    double* Dt = p.tn;    // D_temp
    double* DtA = p.tn2;  // D_temp_A
    double* Et = p.tm;    // E_temp
    S.c = 1.0;
    for (int i = t.tid; i < n; i += t.T) p.D[i] = 1.0;
    for (int i = t.tid; i < m; i += t.T) p.E[i] = 1.0;
    t.sync();
    for (int it = 0; it < p.set.scaling; ++it) {
      // inf-norms of the KKT columns: P columns, A columns (= A' rows), A
      // rows.  A column norm of P is taken over the stored upper-triangular
      // column, as the builtin algebra's OSQPMatrix_col_norm_inf does
      // (algebra/builtin/matrix.c:194-197, csc_math.c:292-308): column r of
      // triu(P) is row r of the full P restricted to columns <= r.
      for (int r = t.tid; r < n; r += t.T) {
        double a = 0.0, b = 0.0;
        for (int k = p.P.rp[r]; k < p.P.rp[r + 1]; ++k)
          if (p.P.ci[k] <= r) a = dmax(a, fabs(p.P.val[k]));
        for (int k = p.At.rp[r]; k < p.At.rp[r + 1]; ++k) b = dmax(b, fabs(p.At.val[k]));
        Dt[r] = a;
        DtA[r] = b;
      }
      for (int r = t.tid; r < m; r += t.T) {
        double a = 0.0;
        for (int k = p.A.rp[r]; k < p.A.rp[r + 1]; ++k) a = dmax(a, fabs(p.A.val[k]));
        Et[r] = a;
      }
      for (int i = t.tid; i < n; i += t.T) {
        double v = Dt[i] > DtA[i] ? Dt[i] : DtA[i];
        Dt[i] = 1.0 / sqrt(limit_elem(v));
      }
      for (int i = t.tid; i < m; i += t.T) Et[i] = 1.0 / sqrt(limit_elem(Et[i]));
      t.sync();
      scale_P(Dt);
      scale_A(Et, Dt);
      for (int i = t.tid; i < n; i += t.T) {
        p.q[i] = p.q[i] * Dt[i];
        p.D[i] = p.D[i] * Dt[i];
      }
      for (int i = t.tid; i < m; i += t.T) p.E[i] = p.E[i] * Et[i];
      t.sync();
      // cost scaling: mean of P's (upper-triangular) column norms vs ||q||
      for (int r = t.tid; r < n; r += t.T) {
        double a = 0.0;
        for (int k = p.P.rp[r]; k < p.P.rp[r + 1]; ++k)
          if (p.P.ci[k] <= r) a = dmax(a, fabs(p.P.val[k]));
        Dt[r] = a;
      }
      t.sync();
      double c_temp = team_reduce<false>(t, n, [&](int i) { return fabs(Dt[i]); });
      c_temp = c_temp / n;
      double inf_q = norm_inf(p.q, n);
      inf_q = limit_scalar(inf_q);
      c_temp = c_temp > inf_q ? c_temp : inf_q;
      c_temp = limit_scalar(c_temp);
      c_temp = 1.0 / c_temp;
      for (int r = t.tid; r < n; r += t.T)
        for (int k = p.P.rp[r]; k < p.P.rp[r + 1]; ++k) p.P.val[k] *= c_temp;
      for (int i = t.tid; i < n; i += t.T) p.q[i] *= c_temp;
      S.c *= c_temp;  // every thread holds the same value
      t.sync();
    }
    S.cinv = 1.0 / S.c;
    for (int i = t.tid; i < n; i += t.T) p.Dinv[i] = 1.0 / p.D[i];
    for (int i = t.tid; i < m; i += t.T) {
      p.Einv[i] = 1.0 / p.E[i];
      p.l[i] = p.l[i] * p.E[i];
      p.u[i] = p.u[i] * p.E[i];
    }
    t.sync();
  }
  __device__ __forceinline__ void no_scaling() {
    S.c = 1.0;
    S.cinv = 1.0;
    for (int i = t.tid; i < p.n; i += t.T) p.D[i] = p.Dinv[i] = 1.0;
    for (int i = t.tid; i < p.m; i += t.T) p.E[i] = p.Einv[i] = 1.0;
    t.sync();
  }

  // ---- rho (src/auxil.c:76-122, osqp_api.c osqp_update_rho) -------------------
  // classify constraints and set rho / 1/rho; returns whether a type changed
  __device__ __forceinline__ int set_rho_vec() {
    S.rho = dmin(dmax(S.rho, kRhoMin), kRhoMax);
    const double infval = kInfty * kMinScaling;
    const double changed = team_reduce<true>(t, p.m, [&](int i) {
      const int old = p.ctype[i];
      int ty;
      if ((p.l[i] < -infval) && (p.u[i] > infval)) ty = -1;
      else if (p.u[i] - p.l[i] < kRhoTol) ty = 1;
      else ty = 0;
      return ty != old ? 1.0 : 0.0;
    });
    for (int i = t.tid; i < p.m; i += t.T) {
      int ty;
      if ((p.l[i] < -infval) && (p.u[i] > infval)) ty = -1;
      else if (p.u[i] - p.l[i] < kRhoTol) ty = 1;
      else ty = 0;
      p.ctype[i] = ty;
      const double r = ty < 0 ? kRhoMin : (ty == 0 ? S.rho : kRhoEqOverIneq * S.rho);
      p.rho[i] = r;
      p.rhoinv[i] = 1.0 / r;
    }
    t.sync();
    return changed > 0.0;
  }
  __device__ __forceinline__ void update_rho(double rho_new) {
    S.rho = dmin(dmax(rho_new, kRhoMin), kRhoMax);
    for (int i = t.tid; i < p.m; i += t.T) {
      const int ty = p.ctype[i];
      const double r = ty < 0 ? kRhoMin : (ty == 0 ? S.rho : kRhoEqOverIneq * S.rho);
      p.rho[i] = r;
      p.rhoinv[i] = 1.0 / r;
    }
    t.sync();
    update_precond();
  }

  // the row groups' columns and values in slot order, and those of the
  // compact long rows and of the slices, from the CSR arrays (values scaled)
  __device__ __forceinline__ void fill_group_values() {
    const BCsr* Ms[3] = {&p.P, &p.A, &p.At};
    for (int k = 0; k < 3; ++k) {
      const BCsr& M = *Ms[k];
      const int* __restrict__ rp = M.rp;
      const int* __restrict__ ci = M.ci;
      const double* __restrict__ val = M.val;
      const int* __restrict__ grow = M.grow;
      const unsigned char* __restrict__ goff = M.goff;
      int* __restrict__ gci = M.gci;
      double* __restrict__ gval = M.gval;
      const int slots = M.ng * kGroupSlots;
      for (int i = t.tid; i < slots; i += t.T) {
        const int r = grow[i];
        const int e = r >= 0 ? rp[r] + goff[i] : 0;
        gci[i] = r >= 0 ? ci[e] : 0;
        gval[i] = r >= 0 ? val[e] : 0.0;
      }
      if (M.lv32) {
        // long row j: entries loff[j] + q, q below its padded length
        unsigned short* __restrict__ lci16 = M.lci16;
        float* __restrict__ lv32 = M.lv32;
        for (int j = 0; j < M.nlong; ++j) {
          const int r = M.longs[j], s = M.loff[j], len = rp[r + 1] - rp[r], pad = M.loff[j + 1] - s;
          for (int q = t.tid; q < pad; q += t.T) {
            lci16[s + q] = q < len ? (unsigned short)ci[rp[r] + q] : (unsigned short)0;
            lv32[s + q] = q < len ? (float)val[rp[r] + q] : 0.f;
          }
        }
      }
      if (M.ev32) {
        // slice s, lane l: entries q < the slice's width at
        // eoff[s] + (q / kEllVec) kWave kEllVec + l kEllVec + q % kEllVec
        unsigned short* __restrict__ eci16 = M.eci16;
        float* __restrict__ ev32 = M.ev32;
        for (int w = t.tid; w < M.nslice * kWave; w += t.T) {
          const int sl = w / kWave, lane = w % kWave;
          const int r = M.erl[2 * w], len = M.erl[2 * w + 1];
          const int b0 = M.eoff[sl], width = (M.eoff[sl + 1] - b0) / kWave;
          for (int q = 0; q < width; ++q) {
            const int at = b0 + (q / kEllVec) * kWave * kEllVec + lane * kEllVec + q % kEllVec;
            eci16[at] = q < len ? (unsigned short)ci[rp[r] + q] : (unsigned short)0;
            ev32[at] = q < len ? (float)val[rp[r] + q] : 0.f;
          }
        }
      }
    }
    t.sync();
  }

  // Jacobi preconditioner: 1 / (sigma + diag(P) + sum_j rho_j A_ji^2)
  __device__ __forceinline__ void update_precond() {
    V* __restrict__ dinv = (V*)p.cdinv;
    const double* __restrict__ rho = p.rho;
    double* __restrict__ prec = p.prec;
    const int* __restrict__ pdiag = p.Pdiag;
    const double* __restrict__ pval = p.P.val;
    const double sigma = p.set.sigma;
    const int n = p.n;
    team_rows<double>(
        t, p.At, [=](double a, int c) { return a * a * rho[c]; }, [=](int r, double s) { prec[r] = s; });
    t.sync();
    for (int i = t.tid; i < n; i += t.T) {
      const double pd = pdiag[i] >= 0 ? pval[pdiag[i]] : 0.0;
      const double d = sigma + pd + prec[i];
      dinv[i] = (V)(1.0 / d);
    }
    t.sync();
  }

  // ---- PCG with FP64 defect correction (one linear solve) ----------------------
  // b = [b1 (n); b2 (m)] in/out: out = [x~ ; A x~]
  __device__ __forceinline__ void linsys_solve(double* __restrict__ b1, double* __restrict__ b2, int admm_iter) {
    const int n = p.n, m = p.m;
    const double sg = p.set.sigma;
    const int cg_max = p.set.cg_max_iter;
    // arena arrays as global pointers (global loads: their waits are not
    // tied to the LDS accesses' counter)
    const gptr<const double> rho = (gptr<const double>)p.rho;
    const gptr<double> x64 = (gptr<double>)p.x64;
    const gptr<double> r64 = (gptr<double>)p.r64;
    const gptr<double> rhs64 = (gptr<double>)p.rhs64;
    const gptr<double> w64 = (gptr<double>)p.w64;
    const gptr<double> ax64 = (gptr<double>)p.ax64;
    const gptr<double> gb1 = (gptr<double>)b1, gb2 = (gptr<double>)b2;
    const BCsr A = p.A, P = p.P, At = p.At;
    const bool guard = sizeof(V) == 4;
    // zero-iteration bookkeeping of the previous solve (cuda_pcg_interface.cu:266-268)
    if (S.cg_solves > 0) {
      if (S.cg_iter == 0) S.cg_zero_iters++;
      else S.cg_zero_iters = 0;
    }
    const bool cached = m > 0 && S.ax_valid && admm_iter > 1;
    double eps;
    if (!cached) {
      // rhs64 = b1 + A'(rho o b2); ||rhs64||_inf
      team_rows<double>(
          t, At, [=](double v, int c) { return v * (rho[c] * gb2[c]); },
          [=](int row, double s) { rhs64[row] = gb1[row] + s; });
      t.sync();
      const double rhs_norm = norm_inf(p.rhs64, n);
      if (admm_iter == 1) {
        S.cg_red_factor = p.set.cg_tol_fraction;
        S.cg_eps_prev = (rhs_norm < kCgTolMin) ? 1.0 : rhs_norm * S.cg_red_factor;
        eps = S.cg_eps_prev;
      } else {
        eps = next_tolerance();
      }
      // w64 = rho o (A x64); r64 = sigma x64 + P x64 + A' w64 - rhs64
      if (m) {
        team_rows<double>(
            t, A, [=](double v, int c) { return v * x64[c]; }, [=](int row, double s) { w64[row] = rho[row] * s; });
        t.sync();
      }
      residual64(p.rhs64);
    } else {
      eps = next_tolerance();
      for (int i = t.tid; i < m; i += t.T) w64[i] = rho[i] * (ax64[i] - gb2[i]);
      t.sync();
      residual64(b1);
    }
    const double rn = norm_inf(p.r64, n);
    S.cg_solves++;
    const double scale = (rn > 0.0 && isfinite(rn)) ? rn : 1.0;
    const double eps_cmp = eps / scale;
    const double inv = 1.0 / scale;
    // CG on K d = -r64 / scale (init fused with r'y and max|r|), over the CG
    // vectors d, p, Kp, r, y, M^-1 (n) and w (m) as LDS or global pointers
    auto cg = [&](auto d, auto pp, auto Kp, auto r, auto y, auto w, auto dinv) -> int {
      double rty, rnorm;
      team_dot_max<V>(
          t, n,
          [=](int i, V& a, double& c) {
            const V ri = (V)(r64[i] * inv);
            const V yi = dinv[i] * ri;
            r[i] = ri;
            y[i] = yi;
            pp[i] = -yi;
            d[i] = (V)0;
            a = yi * ri;
            c = fabs((double)ri);
          },
          &rty, &rnorm);
      int it = 0;
      bool go = rnorm > eps_cmp && it < cg_max;
      if (guard && !(rty > 0.0 && isfinite(rty))) go = false;
      while (go) {
        if (m) {
          team_rows<V, SL>(
              t, A, [=](double v, int c) { return (V)v * pp[c]; }, [=](int row, V s) { w[row] = (V)rho[row] * s; });
          t.sync();
        }
        // Kp = sigma p + P p + A' w
        team_rows<V, SL>(
            t, P, [=](double v, int c) { return (V)v * pp[c]; }, [=](int row, V s) { Kp[row] = s; });
        t.sync();
        if (m)
          team_rows<V, SL>(
              t, At, [=](double v, int c) { return (V)v * w[c]; },
              [=](int row, V s) { Kp[row] = (V)sg * pp[row] + Kp[row] + s; });
        else
          for (int i = t.tid; i < n; i += t.T) Kp[i] = (V)sg * pp[i] + Kp[i];
        t.sync();
        const double pkp = team_dot<V>(t, n, [=](int i) { return pp[i] * Kp[i]; });
        if (guard && !(pkp > 0.0 && isfinite(pkp))) break;
        const V alpha = (V)(rty / pkp);
        double rty_new;
        team_dot_max<V>(
            t, n,
            [=](int i, V& a, double& c) {
              d[i] += alpha * pp[i];
              const V ri = r[i] + alpha * Kp[i];
              const V yi = dinv[i] * ri;
              r[i] = ri;
              y[i] = yi;
              a = yi * ri;
              c = fabs((double)ri);
            },
            &rty_new, &rnorm);
        const double beta = rty_new / rty;
        it++;
        rty = rty_new;
        go = rnorm > eps_cmp && it < cg_max;
        if (guard && !(rty > 0.0 && isfinite(rty))) go = false;
        if (go) {
          const V vb = (V)beta;
          for (int i = t.tid; i < n; i += t.T) pp[i] = vb * pp[i] - y[i];
          t.sync();
        }
      }
      // x64 += scale d; b1 = x64
      for (int i = t.tid; i < n; i += t.T) {
        double xi = x64[i];
        if (it > 0) xi += scale * (double)d[i];
        x64[i] = xi;
        gb1[i] = xi;
      }
      return it;
    };
    // a small problem keeps its CG vectors in the workgroup's dynamic LDS
    // (d, p, Kp, r, y, M^-1 over n; w over m); a larger one in its arena block.
    // Same arithmetic either way.
    int it;
    if ((size_t)(6 * n + m) * sizeof(V) <= (size_t)t.lds_bytes) {
      const lptr<V> lv = (lptr<V>)(V*)t.lds;
      const gptr<const V> gdinv = (gptr<const V>)(const V*)p.cdinv;
      const lptr<V> dl = lv + 5 * n;
      for (int i = t.tid; i < n; i += t.T) dl[i] = gdinv[i];
      t.sync();
      it = cg(lv, lv + n, lv + 2 * n, lv + 3 * n, lv + 4 * n, lv + 6 * n, (lptr<const V>)dl);
    } else {
      it = cg((gptr<V>)(V*)p.cd, (gptr<V>)(V*)p.cp, (gptr<V>)(V*)p.cKp, (gptr<V>)(V*)p.cr, (gptr<V>)(V*)p.cy,
              (gptr<V>)(V*)p.cw, (gptr<const V>)(const V*)p.cdinv);
    }
    S.cg_iter = it;
    S.cg_total += it;
    t.sync();
    // b2 = A x64 (the A x cache)
    if (m) {
      team_rows<double>(
          t, A, [=](double v, int c) { return v * x64[c]; },
          [=](int row, double s) {
            ax64[row] = s;
            gb2[row] = s;
          });
      t.sync();
      S.ax_valid = 1;
    }
  }
  // tolerance of a non-first ADMM iteration (cuda_pcg_interface.cu:46-60)
  __device__ __forceinline__ double next_tolerance() {
    if (S.cg_zero_iters >= p.set.cg_red_thresh) {
      S.cg_red_factor /= 2;
      S.cg_zero_iters = 0;
    }
    double eps = S.cg_red_factor * sqrt(S.scaled_prim_res * S.scaled_dual_res);
    eps = eps < S.cg_eps_prev ? eps : S.cg_eps_prev;
    eps = eps > kCgTolMin ? eps : kCgTolMin;
    S.cg_eps_prev = eps;
    return eps;
  }
  // r64 = sigma x64 + P x64 + A' w64 - rhs
  __device__ __forceinline__ void residual64(const double* __restrict__ rhs) {
    const int n = p.n;
    const double sigma = p.set.sigma;
    const double* __restrict__ x64 = p.x64;
    const double* __restrict__ w64 = p.w64;
    double* __restrict__ tn = p.tn;
    double* __restrict__ r64 = p.r64;
    team_rows<double>(
        t, p.P, [=](double v, int c) { return v * x64[c]; }, [=](int row, double s) { tn[row] = s; });
    t.sync();
    if (p.m)
      team_rows<double>(
          t, p.At, [=](double v, int c) { return v * w64[c]; },
          [=](int row, double s) { r64[row] = sigma * x64[row] + tn[row] + s - rhs[row]; });
    else
      for (int i = t.tid; i < n; i += t.T) r64[i] = sigma * x64[i] + tn[i] - rhs[i];
    t.sync();
  }

  // ---- ADMM pieces (src/auxil.c) ----------------------------------------------
  __device__ __forceinline__ void cold_start() {
    for (int i = t.tid; i < p.n; i += t.T) {
      x()[i] = 0.0;
      p.x64[i] = 0.0;
    }
    for (int i = t.tid; i < p.m; i += t.T) {
      z()[i] = 0.0;
      p.y[i] = 0.0;
    }
    S.ax_valid = 0;
    t.sync();
  }
  // osqp_warm_start (osqp_api.c:1260-1293) from raw x0 / y0
  __device__ __forceinline__ void warm_start(const double* x0, const double* y0) {
    if (x0) {
      for (int i = t.tid; i < p.n; i += t.T) {
        double v = x0[i];
        if (p.set.scaling) v = v * p.Dinv[i];
        x()[i] = v;
      }
    }
    if (y0) {
      for (int i = t.tid; i < p.m; i += t.T) {
        double v = y0[i];
        if (p.set.scaling) v = (v * p.Einv[i]) * S.c;
        p.y[i] = v;
      }
    }
    t.sync();
    if (x0) {
      double* zz = z();
      double* xx = x();
      team_rows<double>(
          t, p.A, [=](double v, int c) { return v * xx[c]; }, [=](int row, double s) { zz[row] = s; });
      t.sync();
    }
    for (int i = t.tid; i < p.n; i += t.T) p.x64[i] = x()[i];
    S.ax_valid = 0;
    t.sync();
  }

  __device__ __forceinline__ void compute_obj_gap(const double* xx, const double* yy) {
    const double quad = dot(p.Px, xx, p.n);
    const double lin = dot(p.q, xx, p.n);
    const double infval = kInfty * kMinScaling;
    // y projected on the polar recession cone, deadzone rounded
    for (int i = t.tid; i < p.m; i += t.T) {
      double v = yy[i];
      if (p.u[i] > +infval) {
        if (p.l[i] < -infval) v = 0.0;
        else v = dmin(v, 0.0);
      } else if (p.l[i] < -infval) {
        v = dmax(v, 0.0);
      }
      if (fabs(v) < kZeroDeadzone) v = 0.0;
      p.tm[i] = v;
    }
    t.sync();
    double sup = team_reduce<false>(t, p.m, [&](int i) { return p.u[i] * dmax(p.tm[i], 0.0); });
    sup += team_reduce<false>(t, p.m, [&](int i) { return p.l[i] * dmin(p.tm[i], 0.0); });
    double prim = 0.5 * quad + lin;
    double dual = -0.5 * quad - sup;
    S.scaled_gap = quad + lin + sup;
    if (p.set.scaling) {
      prim *= S.cinv;
      dual *= S.cinv;
      S.gap = S.cinv * S.scaled_gap;
    } else {
      S.gap = S.scaled_gap;
    }
    S.obj_val = prim;
    S.dual_obj_val = dual;
    S.xtPx = quad;
    S.qtx = lin;
    S.SC = sup;
  }

  __device__ __forceinline__ double compute_prim_res() {
    double* xx = x();
    const double* zz = z();
    team_spmv(t, p.A, xx, p.Ax);
    t.sync();
    for (int i = t.tid; i < p.m; i += t.T) p.tm[i] = p.Ax[i] - zz[i];
    t.sync();
    S.scaled_prim_res = norm_inf(p.tm, p.m);
    if (unscaled_term()) return scaled_norm_inf(p.Einv, p.tm, p.m);
    return S.scaled_prim_res;
  }
  __device__ __forceinline__ double compute_dual_res() {
    double* xx = x();
    team_spmv(t, p.P, xx, p.Px);
    if (p.m) team_spmv(t, p.At, p.y, p.Aty);
    t.sync();
    for (int i = t.tid; i < p.n; i += t.T) {
      double v = p.q[i] + p.Px[i];
      if (p.m) v = v + p.Aty[i];
      p.tn[i] = v;
    }
    t.sync();
    S.scaled_dual_res = norm_inf(p.tn, p.n);
    if (unscaled_term()) return S.cinv * scaled_norm_inf(p.Dinv, p.tn, p.n);
    return S.scaled_dual_res;
  }
  __device__ __forceinline__ void update_info(int iter) {
    (void)iter;
    S.prim_res = p.m == 0 ? 0.0 : compute_prim_res();
    S.dual_res = compute_dual_res();
    compute_obj_gap(x(), p.y);
    S.primdual_int += fabs(S.gap);
    S.rel_kkt = dmax(dmax(S.dual_res, S.prim_res), S.gap);
  }

  __device__ __forceinline__ double prim_tol(double ea, double er) {
    double a, b;
    if (unscaled_term()) {
      a = scaled_norm_inf(p.Einv, z(), p.m);
      b = scaled_norm_inf(p.Einv, p.Ax, p.m);
    } else {
      a = norm_inf(z(), p.m);
      b = norm_inf(p.Ax, p.m);
    }
    return ea + er * dmax(a, b);
  }
  __device__ __forceinline__ double dual_tol(double ea, double er) {
    double r, s;
    if (unscaled_term()) {
      r = scaled_norm_inf(p.Dinv, p.q, p.n);
      s = scaled_norm_inf(p.Dinv, p.Aty, p.n);
      r = dmax(r, s);
      s = scaled_norm_inf(p.Dinv, p.Px, p.n);
      r = dmax(r, s);
      r *= S.cinv;
    } else {
      r = norm_inf(p.q, p.n);
      s = norm_inf(p.Aty, p.n);
      r = dmax(r, s);
      s = norm_inf(p.Px, p.n);
      r = dmax(r, s);
    }
    return ea + er * r;
  }
  __device__ __forceinline__ double gap_tol(double ea, double er) {
    double r = fabs(S.xtPx);
    r = dmax(r, fabs(S.qtx));
    r = dmax(r, fabs(S.SC));
    if (unscaled_term()) r = S.cinv * r;
    return ea + er * r;
  }
  __device__ __forceinline__ bool is_primal_infeasible(double eps_pinf) {
    const double infval = kInfty * kMinScaling;
    for (int i = t.tid; i < p.m; i += t.T) {
      double v = p.dy[i];
      if (p.u[i] > +infval) {
        if (p.l[i] < -infval) v = 0.0;
        else v = dmin(v, 0.0);
      } else if (p.l[i] < -infval) {
        v = dmax(v, 0.0);
      }
      p.dy[i] = v;
    }
    t.sync();
    double ndy;
    if (unscaled_term()) ndy = team_reduce<true>(t, p.m, [&](int i) { return fabs(p.E[i] * p.dy[i]); });
    else ndy = norm_inf(p.dy, p.m);
    if (ndy > kDivisionTol) {
      double lhs = team_reduce<false>(t, p.m, [&](int i) { return p.u[i] * dmax(p.dy[i], 0.0); });
      lhs += team_reduce<false>(t, p.m, [&](int i) { return p.l[i] * dmin(p.dy[i], 0.0); });
      if (lhs < 0.0) {
        team_spmv(t, p.At, p.dy, p.tn);
        t.sync();
        double nat;
        if (unscaled_term()) nat = team_reduce<true>(t, p.n, [&](int i) { return fabs(p.tn[i] * p.Dinv[i]); });
        else nat = norm_inf(p.tn, p.n);
        return nat < eps_pinf * ndy;
      }
    }
    return false;
  }
  __device__ __forceinline__ bool is_dual_infeasible(double eps_dinf) {
    double ndx, cs;
    if (unscaled_term()) {
      ndx = scaled_norm_inf(p.D, p.dx, p.n);
      cs = S.c;
    } else {
      ndx = norm_inf(p.dx, p.n);
      cs = 1.0;
    }
    if (ndx > kDivisionTol) {
      if (dot(p.q, p.dx, p.n) < 0.0) {
        team_spmv(t, p.P, p.dx, p.tn);
        t.sync();
        double npd;
        if (unscaled_term()) npd = team_reduce<true>(t, p.n, [&](int i) { return fabs(p.tn[i] * p.Dinv[i]); });
        else npd = norm_inf(p.tn, p.n);
        if (npd < cs * eps_dinf * ndx) {
          team_spmv(t, p.A, p.dx, p.tm);
          t.sync();
          if (unscaled_term())
            for (int i = t.tid; i < p.m; i += t.T) p.tm[i] = p.tm[i] * p.Einv[i];
          t.sync();
          const double infval = kInfty * kMinScaling, tol = eps_dinf * ndx;
          const double bad = team_reduce<true>(t, p.m, [&](int i) {
            const double yv = p.tm[i], l = p.l[i], u = p.u[i];
            return (((u < +infval) && (yv > +tol)) || ((l > -infval) && (yv < -tol))) ? 1.0 : 0.0;
          });
          return bad == 0.0;
        }
      }
    }
    return false;
  }
  // returns 1 when the solve ends (src/auxil.c:808-945)
  __device__ __forceinline__ int check_termination(int approximate) {
    double ea = p.set.eps_abs, er = p.set.eps_rel, epi = p.set.eps_pinf, edi = p.set.eps_dinf;
    if ((S.prim_res > kInfty) || (S.dual_res > kInfty)) {
      S.status = kNonCvx;
      S.obj_val = kOsqpNan;
      return 1;
    }
    if (approximate) {
      ea *= 10;
      er *= 10;
      epi *= 10;
      edi *= 10;
    }
    bool prim_ok = false, dual_ok = false, gap_ok = false, prim_inf = false, dual_inf = false;
    if (p.m == 0) {
      prim_ok = true;
    } else {
      if (S.prim_res < prim_tol(ea, er)) prim_ok = true;
      else prim_inf = is_primal_infeasible(epi);
    }
    if (S.dual_res < dual_tol(ea, er)) dual_ok = true;
    else dual_inf = is_dual_infeasible(edi);
    if (p.set.check_gap) {
      const double eg = gap_tol(ea, er);
      if (unscaled_term()) gap_ok = fabs(S.gap) < eg;
      else gap_ok = fabs(S.scaled_gap) < eg;
    } else {
      gap_ok = true;
    }
    if (prim_ok && dual_ok && gap_ok) {
      S.status = approximate ? kSolvedInaccurate : kSolved;
      return 1;
    } else if (prim_inf) {
      S.status = approximate ? kPrimalInfeasibleInaccurate : kPrimalInfeasible;
      if (unscaled_term()) {
        for (int i = t.tid; i < p.m; i += t.T) p.dy[i] = p.dy[i] * p.E[i];
        t.sync();
      }
      S.obj_val = kInfty;
      return 1;
    } else if (dual_inf) {
      S.status = approximate ? kDualInfeasibleInaccurate : kDualInfeasible;
      if (unscaled_term()) {
        for (int i = t.tid; i < p.n; i += t.T) p.dx[i] = p.dx[i] * p.D[i];
        t.sync();
      }
      S.obj_val = -kInfty;
      return 1;
    }
    return 0;
  }

  __device__ __forceinline__ double rho_estimate() {
    const double pr = S.scaled_prim_res / (dmax(norm_inf(z(), p.m), norm_inf(p.Ax, p.m)) + kDivisionTol);
    double dn = dmax(norm_inf(p.q, p.n), norm_inf(p.Aty, p.n));
    dn = dmax(dn, norm_inf(p.Px, p.n));
    const double du = S.scaled_dual_res / (dn + kDivisionTol);
    const double est = S.rho * sqrt(pr / du);
    return dmin(dmax(est, kRhoMin), kRhoMax);
  }

  __device__ __forceinline__ bool has_solution() const {
    const int s = S.status;
    return s != kPrimalInfeasible && s != kPrimalInfeasibleInaccurate && s != kDualInfeasible &&
           s != kDualInfeasibleInaccurate && s != kNonCvx;
  }

  // setup of a new problem (osqp_setup: scale_data, set_rho_vec, linsys
  // init, cold start) and an optional osqp_warm_start
  __device__ __forceinline__ void setup() {
    S.t_start = (long long)wall_clock64();
    const int4* s4 = (const int4*)p.stage;
    int4* d4 = (int4*)p.arena_static;
    const long long n4 = p.stage_bytes / 16;
    for (long long i = t.tid; i < n4; i += t.T) d4[i] = s4[i];
    for (int i = t.tid; i < p.m; i += t.T) p.ctype[i] = 0;
    t.sync();
    S.rho = p.set.rho0;
    if (p.set.scaling) scale_data();
    else no_scaling();
    fill_group_values();
    set_rho_vec();
    update_precond();
    S.swapped = 0;
    cold_start();
    if (p.warm_x || p.warm_y) warm_start(p.warm_x, p.warm_y);
    S.status = kUnsolved;
    S.iter = 0;
    S.rho_updates = 0;
    S.can_check = 0;
    S.obj_val = kInfty;
    S.prim_res = kInfty;
    S.dual_res = kInfty;
    S.rel_kkt = kInfty;
    S.last_rel_kkt = kInfty;
    S.rho_estimate = S.rho;
    S.primdual_int = 0.0;
    S.cg_iter = 0;
    S.cg_zero_iters = 0;
    S.cg_solves = 0;
    S.cg_total = 0;
    S.cg_eps_prev = 0.0;
    S.cg_red_factor = p.set.cg_tol_fraction;
    S.ad_interval = p.set.ad_interval;
    S.t_solve = (long long)wall_clock64();
    S.setup_ticks = S.t_solve - S.t_start;
    S.phase = kPhaseRun;
  }

  // osqp_update_data_mat (src/osqp_api.c:1309-1409) on the GPU: the data
  // are unscaled (src/scaling.c:169-192, same operation order), the new raw
  // values written, the data scaled again from scratch and the
  // preconditioner rebuilt; the iterates stay as they are
  __device__ __forceinline__ void update_matrices(const BMatUpd& u) {
    const int n = p.n, m = p.m;
    if (p.set.scaling) {
      const double ci = S.cinv;
      for (int r = t.tid; r < n; r += t.T)
        for (int k = p.P.rp[r]; k < p.P.rp[r + 1]; ++k) {
          const int c = p.P.ci[k];
          const int lo = c < r ? c : r, hi = c < r ? r : c;
          p.P.val[k] = ((p.P.val[k] * ci) * p.Dinv[lo]) * p.Dinv[hi];
        }
      for (int r = t.tid; r < m; r += t.T)
        for (int k = p.A.rp[r]; k < p.A.rp[r + 1]; ++k) p.A.val[k] = (p.A.val[k] * p.Einv[r]) * p.Dinv[p.A.ci[k]];
      for (int r = t.tid; r < n; r += t.T) {
        for (int k = p.At.rp[r]; k < p.At.rp[r + 1]; ++k)
          p.At.val[k] = (p.At.val[k] * p.Einv[p.At.ci[k]]) * p.Dinv[r];
        p.q[r] = (p.q[r] * ci) * p.Dinv[r];
      }
      for (int i = t.tid; i < m; i += t.T) {
        p.l[i] = p.l[i] * p.Einv[i];
        p.u[i] = p.u[i] * p.Einv[i];
      }
      t.sync();
    }
    for (int k = t.tid; k < u.np; k += t.T) p.P.val[u.pslot[k]] = u.pval[k];
    for (int k = t.tid; k < u.na; k += t.T) {
      p.A.val[u.aslot[k]] = u.aval[k];
      p.At.val[u.tslot[k]] = u.aval[k];
    }
    t.sync();
    if (p.set.scaling) scale_data();
    fill_group_values();
    update_precond();
    S.ax_valid = 0;
    t.sync();
  }

  // a new solve of a resident problem: osqp_update_data_vec(q) (if given),
  // reset_info, and osqp_solve's cold start when warm starting is off
  __device__ __forceinline__ void new_solve(const double* q_new) {
    S.t_start = (long long)wall_clock64();
    if (q_new) {
      for (int i = t.tid; i < p.n; i += t.T) {
        double v = q_new[i];
        if (p.set.scaling) v = (v * p.D[i]) * S.c;
        p.q[i] = v;
      }
      t.sync();
    }
    S.status = kUnsolved;
    S.rho_updates = 0;
    S.iter = 0;
    S.can_check = 0;
    if (!p.set.warm) cold_start();
    S.t_solve = (long long)wall_clock64();
    S.phase = kPhaseRun;
  }

  // end of a solve: objective, final status, unscaled solution and
  // certificates to the host (src/osqp_api.c:908-948, src/auxil.c:598-674)
  __device__ __forceinline__ void finish() {
    int iter_out = S.iter;
    if (!S.can_check) {
      update_info(S.iter);
      check_termination(0);
    }
    // a time-limit stop reports the last completed iteration: the loop
    // variable's value less one, as update_info(iter - 1) after the loop does
    if (S.status == kTimeLimit) iter_out = S.iter - 1;
    if (has_solution()) compute_obj_gap(x(), p.y);
    if (S.status == kUnsolved) {
      if (!check_termination(1)) S.status = kMaxIterReached;
    }
    if (S.status == kTimeLimit) {
      if (!check_termination(1)) S.status = kTimeLimit;
    }
    S.rho_estimate = rho_estimate();
    const bool prim_inf = S.status == kPrimalInfeasible || S.status == kPrimalInfeasibleInaccurate;
    const bool dual_inf = S.status == kDualInfeasible || S.status == kDualInfeasibleInaccurate;
    if (has_solution()) {
      for (int i = t.tid; i < p.n; i += t.T) {
        p.out_x[i] = p.set.scaling ? x()[i] * p.D[i] : x()[i];
        p.out_dcert[i] = kOsqpNan;
      }
      for (int i = t.tid; i < p.m; i += t.T) {
        p.out_y[i] = p.set.scaling ? (p.y[i] * p.E[i]) * S.cinv : p.y[i];
        p.out_pcert[i] = kOsqpNan;
      }
    } else {
      for (int i = t.tid; i < p.n; i += t.T) p.out_x[i] = kOsqpNan;
      for (int i = t.tid; i < p.m; i += t.T) p.out_y[i] = kOsqpNan;
      // the certificate normalised by its inf-norm (mult_scalar by 1 / norm);
      // the other one is OSQP_NAN; a non-convex stop writes neither
      if (prim_inf) {
        const double inv = 1.0 / norm_inf(p.dy, p.m);
        for (int i = t.tid; i < p.m; i += t.T) p.out_pcert[i] = p.dy[i] * inv;
        for (int i = t.tid; i < p.n; i += t.T) p.out_dcert[i] = kOsqpNan;
      }
      if (dual_inf) {
        const double inv = 1.0 / norm_inf(p.dx, p.n);
        for (int i = t.tid; i < p.n; i += t.T) p.out_dcert[i] = p.dx[i] * inv;
        for (int i = t.tid; i < p.m; i += t.T) p.out_pcert[i] = kOsqpNan;
      }
      t.sync();
      cold_start();
    }
    t.sync();
    if (t.tid == 0) {
      BResult* o = p.out;
      o->status = S.status;
      o->iter = iter_out;
      o->rho_updates = S.rho_updates;
      o->obj_val = S.obj_val;
      o->dual_obj_val = S.dual_obj_val;
      o->prim_res = S.prim_res;
      o->dual_res = S.dual_res;
      o->gap = S.gap;
      o->rho_estimate = S.rho_estimate;
      o->primdual_int = S.primdual_int;
      o->rel_kkt = S.rel_kkt;
      o->cg_iters = S.cg_total;
      __threadfence_system();
      o->done = 1;
    }
    S.phase = kPhaseDone;
  }

  // one ADMM iteration; returns 1 when the solve has ended
  __device__ __forceinline__ int iterate() {
    const int iter = ++S.iter;
    const double a = p.set.alpha, sigma = p.set.sigma;
    const int n = p.n, m = p.m;
    S.swapped ^= 1;  // swap(x, x_prev), swap(z, z_prev)
    double* __restrict__ xx = x();
    const double* __restrict__ xp = xprev();
    double* __restrict__ zz = z();
    const double* __restrict__ zp = zprev();
    double* __restrict__ xt = p.xt;
    double* __restrict__ zt = p.zt;
    double* __restrict__ y = p.y;
    double* __restrict__ dx = p.dx;
    double* __restrict__ dy = p.dy;
    const double* __restrict__ q = p.q;
    const double* __restrict__ rho = p.rho;
    const double* __restrict__ rhoinv = p.rhoinv;
    const double* __restrict__ l = p.l;
    const double* __restrict__ u = p.u;
    // compute_rhs
    for (int i = t.tid; i < n; i += t.T) xt[i] = sigma * xp[i] + (-1.0) * q[i];
    for (int i = t.tid; i < m; i += t.T) {
      const double v = rhoinv[i] * y[i];
      zt[i] = (-1.0) * v + 1.0 * zp[i];
    }
    t.sync();
    linsys_solve(xt, zt, iter);
    // update_x, update_z, update_y
    for (int i = t.tid; i < n; i += t.T) {
      const double v = a * xt[i] + (1.0 - a) * xp[i];
      xx[i] = v;
      dx[i] = v - xp[i];
    }
    for (int i = t.tid; i < m; i += t.T) {
      double v = rhoinv[i] * y[i];
      v += a * zt[i] + (1.0 - a) * zp[i];  // add_scaled3's increment form (x == a, sca == 1)
      v = dmin(dmax(v, l[i]), u[i]);
      zz[i] = v;
      double dyv = a * zt[i] + (1.0 - a) * zp[i] + (-1.0) * v;
      dyv = dyv * rho[i];
      dy[i] = dyv;
      y[i] = y[i] + dyv;
    }
    t.sync();
    // time limit (src/osqp_api.c:743-759), before this iteration's checks
    if (p.set.tl_ticks > 0 && t.decide((long long)wall_clock64() - S.t_start >= p.set.tl_ticks)) {
      S.status = kTimeLimit;
      return 1;
    }
    const int ct = p.set.check_term;
    const bool can_check = ct && (iter % ct == 0);
    bool can_adapt = false;
    switch (p.set.adaptive) {
      case kAdaptTime:
        // src/osqp_api.c:782-820: once the loop has run for the fraction of
        // the setup time, the interval is fixed from the iteration count
        // (with the reference's min(1, .) bound)
        if (S.ad_interval == 0) {
          if (!t.decide((double)((long long)wall_clock64() - S.t_solve) >
                        p.set.ad_fraction * (double)S.setup_ticks))
            break;
          const double base = ct ? ct : kCheckTermDefault;
          const double xv = (double)iter + 0.5 * base;
          int iv = (int)(xv - fmod(xv, base));
          iv = iv > ct ? iv : ct;
          S.ad_interval = iv < 1 ? iv : 1;
        }
        // fall through
      case kAdaptIters:
      case kAdaptKkt:
        can_adapt = S.ad_interval && (iter % S.ad_interval == 0);
        break;
      default:
        break;
    }
    S.can_check = can_check;
    if (can_check || can_adapt || iter == 1) update_info(iter);
    if (can_check && check_termination(0)) return 1;
    // src/osqp_api.c:868-877: the KKT mode adapts only after the KKT error
    // has fallen by the fraction since the last update
    if (can_adapt && p.set.adaptive == kAdaptKkt) can_adapt = S.rel_kkt <= p.set.ad_fraction * S.last_rel_kkt;
    if (can_adapt) {
      const double rho_new = rho_estimate();
      S.rho_estimate = rho_new;
      if ((rho_new > S.rho * p.set.ad_tol) || (rho_new < S.rho / p.set.ad_tol)) {
        update_rho(rho_new);
        S.rho_updates += 1;
        S.last_rel_kkt = S.rel_kkt;
      }
    }
    return iter >= p.set.max_iter ? 1 : 0;
  }
};

}  // namespace osqp_serve
