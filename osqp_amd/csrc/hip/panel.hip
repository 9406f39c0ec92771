// panel.hip — column-panelled SpMV: host construction and the gfx950 kernel.
//
// Construction is integer-exact: entries keep their row order and, within a
// row, their column order; panel p holds the columns [p W, (p + 1) W).  `src`
// records, for every panel entry, the CSR slot it came from, so values are
// refreshed on the device with one gather whenever the source matrix is
// rescaled or updated.
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "panel.h"

namespace osqp_hip {

namespace {

constexpr int kPanelWaves = 16;  // wavefronts per workgroup (1024 threads)

__global__ void k_panel_gather(long long nnz, const int* __restrict__ src, const double* __restrict__ csr_val,
                               double* __restrict__ v64, float* __restrict__ v32) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (long long)gridDim.x * blockDim.x) {
    const double v = csr_val[src[k]];
    if (v64) v64[k] = v;
    if (v32) v32[k] = (float)v;
  }
}

template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) {
  return __builtin_nontemporal_load(p);
}

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// Entries per lane per load group: a lane loads V consecutive (index, value)
// pairs with one 16-byte value load and one 8- or 4-byte index load (the
// address unit's cost is per instruction, so wide loads move more per slot).
template <typename TV>
constexpr int panel_vec() { return sizeof(TV) == 4 ? 4 : 2; }

// One task per wavefront at a time; see panel.h for the scheme.  V is the
// vector width the panels were built for (their task cap); FP32 values on
// FP64-built panels (V = 2) are loaded as 8-byte pairs and multiplied and
// summed in TX.
template <typename TV, typename TX, int W, int NW, int PER, int V, bool SEL>
__global__ __launch_bounds__(NW * 64) void k_panel_spmv(const int* active, int ntask, const int4* __restrict__ task,
                                                        const int2* __restrict__ tpanel, const int* __restrict__ rp,
                                                        const uint16_t* __restrict__ ci, const TV* __restrict__ val,
                                                        int R, int C, const int* __restrict__ pcol,
                                                        const TX* __restrict__ x0, int n0,
                                                        const TX* __restrict__ x1, TX* __restrict__ partial) {
  if (active && !*active) return;
  static_assert(V == 4 ? (sizeof(TV) == 4 && sizeof(TX) == 4) : V == 2, "vector widths: 4 (FP32) or 2");
  constexpr int NG = PER / V;          // load groups of 64 V entries
  constexpr int CAP = 64 * PER - (V - 1);  // entries of a task (the aligned window holds 64 PER)
  static_assert(PER % V == 0, "PER must be a multiple of the vector width");
  __shared__ __attribute__((aligned(16))) TX xs[W];
  __shared__ __attribute__((aligned(16))) TX prod_all[NW][64 * PER];
  constexpr int XL = W / (NW * 64);
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  TX* prod = prod_all[wv];
  const int t_begin = (int)(((long long)blockIdx.x * ntask) / gridDim.x);
  const int t_end = (int)(((long long)(blockIdx.x + 1) * ntask) / gridDim.x);
  for (int t = t_begin; t < t_end;) {
    const int2 pe = tpanel[t];
    const int p = pe.x;
    const int te = pe.y < t_end ? pe.y : t_end;
    // the panel's slice of x into LDS: every load in flight, then the stores.
    // A full panel lying wholly in x0 or wholly in x1, 16-byte aligned there,
    // goes as 16-byte vectors (thread i takes vectors i, i + 1024, ...);
    // otherwise element by element.  The choice is the same for the whole
    // workgroup.
    const int c0 = (SEL ? pcol[p] : p) * W;  // SEL: panels of selected column blocks only
    const int w = (C - c0) < W ? (C - c0) : W;
    constexpr int XV = 16 / (int)sizeof(TX);  // elements per vector
    constexpr int XG = XL / XV;               // vectors per thread
    static_assert(XL % XV == 0, "the slice must split into whole vectors per thread");
    const TX* src = nullptr;
    if (w == W) {
      if (c0 + W <= n0) src = x0 + c0;
      else if (c0 >= n0 && x1) src = x1 + (c0 - n0);
    }
    if (src && ((uintptr_t)src & 15) == 0) {
      typedef TX VT __attribute__((ext_vector_type(XV)));
      VT xv[XG];
#pragma unroll
      for (int g = 0; g < XG; ++g) xv[g] = *(const VT*)(src + (g * NW * 64 + threadIdx.x) * XV);
      __syncthreads();  // the previous panel's readers are done
#pragma unroll
      for (int g = 0; g < XG; ++g) *(VT*)(xs + (g * NW * 64 + threadIdx.x) * XV) = xv[g];
      __syncthreads();
    } else {
      TX xr[XL];
#pragma unroll
      for (int j = 0; j < XL; ++j) {
        const int i = j * NW * 64 + threadIdx.x;
        const int c = c0 + i;
        xr[j] = i < w ? (c < n0 ? x0[c] : x1[c - n0]) : (TX)0;
      }
      __syncthreads();  // the previous panel's readers are done
#pragma unroll
      for (int j = 0; j < XL; ++j) xs[j * NW * 64 + threadIdx.x] = xr[j];
      __syncthreads();
    }
    const int* rpp = rp + (long long)p * (R + 1);
    TX* out = partial + (long long)p * R;
    // this wavefront's tasks of the panel: t + wv + i NW; lane i holds the
    // descriptor of the i-th one (64 at a time)
    for (int ub = t + wv; ub < te; ub += 64 * NW) {
      const int mine = (te - ub + NW - 1) / NW < 64 ? (te - ub + NW - 1) / NW : 64;
      int4 d = make_int4(0, 0, 0, 0);
      if (lane < mine) d = task[ub + lane * NW];
      // wait for the descriptors here, once: otherwise the wait at the head of
      // the task loop also waits for the previous task's partial-sum stores
      asm volatile("" : "+v"(d.x), "+v"(d.y), "+v"(d.z), "+v"(d.w));
      for (int i = 0; i < mine; ++i) {
        const int rb = __builtin_amdgcn_readlane(d.x, i);
        const int re = __builtin_amdgcn_readlane(d.y, i);
        const int k0 = __builtin_amdgcn_readlane(d.z, i);
        const int len = __builtin_amdgcn_readlane(d.w, i) - k0;
        if (len > CAP) {
          // one panel row longer than a task: lane-strided passes, then a
          // fixed butterfly
          TX acc = 0;
          for (int b = 0; b < len; b += 64 * PER) {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
              const int e = b + j * 64 + lane;
              if (e < len) acc += (TX)ld_nt(val + k0 + e) * xs[ld_nt(ci + k0 + e)];
            }
          }
          acc = wave_sum(acc);
          if (lane == 0) out[rb] = acc;
          continue;
        }
        const int rpl = lane <= re - rb ? rpp[rb + lane] : 0;
        // The window of 64 PER entries from base = k0 rounded down to V holds
        // the task (tasks hold at most CAP entries).  Lane l loads entries
        // base + g 64 V + l V + [0, V) of group g as vectors; every index
        // loaded, in the task or not, is a panel-local column (< W; zero in
        // the padding), so the gathers need no guard; window positions outside
        // [off, off + len) are then stored as zero products.
        const int base = k0 & ~(V - 1), off = k0 - base;
        // all loads first (packed indices: two per 32-bit word), then per
        // group: gathers, products, one 16-byte LDS store
        constexpr int CW = V / 2;  // 32-bit index words per lane per group
        unsigned cw[NG][CW];
        TV vv[NG][V];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          const long long at = (long long)base + g * 64 * V + lane * V;
          if constexpr (V == 4) {
            const u32x2 w2 = ld_nt((const u32x2*)(ci + at));
            const f32x4 v4 = ld_nt((const f32x4*)(val + at));
            cw[g][0] = w2.x;
            cw[g][1] = w2.y;
            vv[g][0] = v4.x;
            vv[g][1] = v4.y;
            vv[g][2] = v4.z;
            vv[g][3] = v4.w;
          } else if constexpr (sizeof(TV) == 8) {
            const unsigned w1 = ld_nt((const unsigned*)(ci + at));
            const f64x2 v2 = ld_nt((const f64x2*)(val + at));
            cw[g][0] = w1;
            vv[g][0] = v2.x;
            vv[g][1] = v2.y;
          } else {
            const unsigned w1 = ld_nt((const unsigned*)(ci + at));
            const f32x2 v2 = ld_nt((const f32x2*)(val + at));
            cw[g][0] = w1;
            vv[g][0] = v2.x;
            vv[g][1] = v2.y;
          }
        }
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          const int q0 = g * 64 * V + lane * V;
          TX pr[V];
#pragma unroll
          for (int i = 0; i < V; ++i) {
            const unsigned col = (cw[g][i >> 1] >> (16 * (i & 1))) & 0xffffu;
            const TX x = xs[col];
            const int q = q0 + i;
            pr[i] = (q >= off && q < off + len) ? (TX)vv[g][i] * x : (TX)0;
          }
          if constexpr (V == 4) {
            f32x4 o;
            o.x = pr[0];
            o.y = pr[1];
            o.z = pr[2];
            o.w = pr[3];
            *(f32x4*)(prod + q0) = o;
          } else {
            f64x2 o;
            o.x = pr[0];
            o.y = pr[1];
            *(f64x2*)(prod + q0) = o;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // L = lanes per row: the largest power of two with L * nr <= 64
        // (64 for nr <= 1); G = 64 / L row groups; all as shifts
        const int nr = re - rb;
        const int lgL = nr <= 1 ? 6 : 6 - (32 - __builtin_clz((unsigned)(nr - 1)));
        const int L = 1 << lgL, lgG = 6 - lgL;
        const int G = 1 << lgG, g = lane >> lgL, gl = lane & (L - 1);
        const int bound = ((nr + G - 1) >> lgG) << lgG;
        for (int r = g; r < bound; r += G) {
          const int s = __shfl(rpl, r & 63) - base;
          const int e = __shfl(rpl, (r + 1) & 63) - base;
          TX acc = 0;
          if (r < nr) {
            int k = s + gl;
            for (; k + 3 * L < e; k += 4 * L) {
              const TX a0 = prod[k], a1 = prod[k + L], a2 = prod[k + 2 * L], a3 = prod[k + 3 * L];
              acc += a0;
              acc += a1;
              acc += a2;
              acc += a3;
            }
            for (; k < e; k += L) acc += prod[k];
          }
          for (int off = L >> 1; off > 0; off >>= 1) acc += __shfl_down(acc, off, L);
          if (r < nr && gl == 0) out[rb + r] = acc;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
    t = te;
  }
}

template <typename TV, typename TX, int W, int PER, int V = panel_vec<TV>()>
void launch(const PanelMatrix& P, const int* active, const TV* val, const TX* x0, int n0, const TX* x1,
            TX* partial) {
  if (P.pcol)
    hipLaunchKernelGGL((k_panel_spmv<TV, TX, W, kPanelWaves, PER, V, true>), dim3(P.grid), dim3(kPanelWaves * 64),
                       0, rt().stream, active, P.ntask, (const int4*)P.task, (const int2*)P.tpanel, (const int*)P.rp,
                       (const uint16_t*)P.ci, val, P.R, P.C, (const int*)P.pcol, x0, n0, x1, partial);
  else
    hipLaunchKernelGGL((k_panel_spmv<TV, TX, W, kPanelWaves, PER, V, false>), dim3(P.grid), dim3(kPanelWaves * 64),
                       0, rt().stream, active, P.ntask, (const int4*)P.task, (const int2*)P.tpanel, (const int*)P.rp,
                       (const uint16_t*)P.ci, val, P.R, P.C, (const int*)nullptr, x0, n0, x1, partial);
}

// FP32 operators: 8192-column panels, 12 nonzeros per lane (80 KB of LDS per
// workgroup); FP64 operators: 4096-column panels, 6 per lane (80 KB).  Two
// workgroups fit a CU.  Measured on the 500000 x 50000 operator with 100
// nonzeros per row: FP32 per = 8 / 10 ran 2-4 % slower than 12, FP64 per = 6
// ran about 4 % faster than 4.
constexpr int kW32 = 8192, kPer32 = 12, kW64 = 4096, kPer64 = 6;

}  // namespace

int panel_width(int bits) { return bits == 32 ? kW32 : kW64; }

int panel_build(PanelMatrix* P, int R, int C, int bits, const int* rp, const int* ci, const unsigned char* sel) {
  memset(P, 0, sizeof(*P));
  if (bits != 32 && bits != 64) return 1;
  const int W = bits == 32 ? kW32 : kW64;
  const int per = bits == 32 ? kPer32 : kPer64;
  const int want32 = bits == 32, want64 = bits == 64;
  const int nblk = C > 0 ? (C + W - 1) / W : 1;
  // column block -> panel (-1: not selected; its entries are skipped)
  std::vector<int> pan_of(nblk), pcol;
  int np = 0;
  for (int b = 0; b < nblk; ++b) {
    pan_of[b] = (!sel || sel[b]) ? np++ : -1;
    if (pan_of[b] >= 0) pcol.push_back(b);
  }
  if (np == 0) return 1;
  const long long pad = 64LL * per;
  P->R = R;
  P->C = C;
  P->W = W;
  P->per = per;
  P->np = np;
  // per (panel, row) counts -> absolute row pointers, panel-major
  std::vector<int> rpa((size_t)np * (R + 1), 0);
  long long nnz = 0;
  for (int r = 0; r < R; ++r)
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      const int pp = pan_of[ci[k] / W];
      if (pp >= 0) {
        rpa[(size_t)pp * (R + 1) + r + 1]++;
        nnz++;
      }
    }
  P->nnz = nnz;
  long long base = 0;
  for (int p = 0; p < np; ++p) {
    int* q = rpa.data() + (size_t)p * (R + 1);
    q[0] = (int)base;
    for (int r = 0; r < R; ++r) q[r + 1] += q[r];
    base = q[R];
  }
  std::vector<uint16_t> lci((size_t)(nnz + pad), 0);
  std::vector<int> src((size_t)(nnz > 0 ? nnz : 1), 0);
  std::vector<int> next(rpa);
  for (int r = 0; r < R; ++r)
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      const int b = ci[k] / W, pp = pan_of[b];
      if (pp < 0) continue;
      const int pos = next[(size_t)pp * (R + 1) + r]++;
      lci[(size_t)pos] = (uint16_t)(ci[k] - b * W);
      src[(size_t)pos] = k;
    }
  // tasks: per panel, greedy row ranges with <= cap nonzeros and <= 63 rows;
  // a panel row longer than cap is a task of its own.  cap leaves room for
  // the kernel's window alignment (it loads 64 per entries from the task's
  // first nonzero rounded down to the vector width).
  const int cap = 64 * per - ((bits == 32 ? 4 : 2) - 1);
  std::vector<int4> tasks;
  std::vector<int2> tp;
  for (int p = 0; p < np && R > 0; ++p) {
    const int* q = rpa.data() + (size_t)p * (R + 1);
    const size_t first = tasks.size();
    int cur = 0;
    for (int i = 0; i < R; ++i) {
      if (q[i + 1] - q[i] > cap) {
        if (i > cur) tasks.push_back(make_int4(cur, i, q[cur], q[i]));
        tasks.push_back(make_int4(i, i + 1, q[i], q[i + 1]));
        cur = i + 1;
      } else if (i > cur && (q[i + 1] - q[cur] > cap || i - cur >= kPanelRowCap)) {
        tasks.push_back(make_int4(cur, i, q[cur], q[i]));
        cur = i;
      }
    }
    if (cur < R) tasks.push_back(make_int4(cur, R, q[cur], q[R]));
    for (size_t k = first; k < tasks.size(); ++k) tp.push_back(make_int2(p, (int)tasks.size()));
  }
  P->ntask = (int)tasks.size();
  P->rp = (int*)dev_alloc(sizeof(int) * rpa.size());
  P->ci = (uint16_t*)dev_alloc(sizeof(uint16_t) * lci.size());
  P->src = (int*)dev_alloc(sizeof(int) * src.size());
  P->task = (int4*)dev_alloc(sizeof(int4) * (tasks.empty() ? 1 : tasks.size()));
  P->tpanel = (int2*)dev_alloc(sizeof(int2) * (tp.empty() ? 1 : tp.size()));
  P->val64 = want64 ? (double*)dev_alloc(sizeof(double) * (size_t)(nnz + pad)) : nullptr;
  P->val32 = want32 ? (float*)dev_alloc(sizeof(float) * (size_t)(nnz + pad)) : nullptr;
  if (!P->rp || !P->ci || !P->src || !P->task || !P->tpanel || (want64 && !P->val64) || (want32 && !P->val32))
    return 1;
  if (np < nblk) {
    if (!(P->pcol = (int*)dev_alloc(sizeof(int) * (size_t)np))) return 1;
    upload(P->pcol, pcol.data(), sizeof(int) * (size_t)np);
  }
  upload(P->rp, rpa.data(), sizeof(int) * rpa.size());
  upload(P->ci, lci.data(), sizeof(uint16_t) * lci.size());
  upload(P->src, src.data(), sizeof(int) * src.size());
  if (!tasks.empty()) {
    upload(P->task, tasks.data(), sizeof(int4) * tasks.size());
    upload(P->tpanel, tp.data(), sizeof(int2) * tp.size());
  }
  // the padding past nnz is read (and multiplied by zero): keep it finite
  if (want64) HIPCHK(hipMemsetAsync(P->val64 + nnz, 0, sizeof(double) * (size_t)pad, rt().stream));
  if (want32) HIPCHK(hipMemsetAsync(P->val32 + nnz, 0, sizeof(float) * (size_t)pad, rt().stream));
  // two resident workgroups per CU, a few rounds of them: 3 rounds for FP32
  // (its tasks are shorter), 4 for FP64 -- the fastest of 1..4 rounds measured
  // on the 500000 x 50000 operator and on [P A'] of the same QP.  An operator
  // with fewer tasks gets about one task per wavefront (a workgroup's x-slice
  // staging then serves 16 tasks, not a handful), but at least one workgroup
  // per CU.
  const int cus = rt().num_cu > 0 ? rt().num_cu : 256;
  long long g = (bits == 32 ? 6LL : 8LL) * cus;
  const long long per_wave = P->ntask / kPanelWaves;
  if (per_wave < g) g = per_wave > cus ? per_wave : cus;
  if (g > P->ntask) g = P->ntask > 0 ? P->ntask : 1;
  P->grid = (int)g;
  P->ver = -1;
  return 0;
}

void panel_free(PanelMatrix* P) {
  dev_free(P->rp);
  dev_free(P->ci);
  dev_free(P->src);
  dev_free(P->task);
  dev_free(P->tpanel);
  dev_free(P->val64);
  dev_free(P->val32);
  dev_free(P->pcol);
  memset(P, 0, sizeof(*P));
}

void panel_refresh_values(PanelMatrix* P, const double* csr_val) {
  if (P->nnz <= 0) return;
  const int g = grid_for(P->nnz, kBlock, 16384);
  hipLaunchKernelGGL(k_panel_gather, dim3(g), dim3(kBlock), 0, rt().stream, P->nnz, (const int*)P->src, csr_val,
                     P->val64, P->val32);
  LAUNCHCHK();
}

void panel_spmv(const PanelMatrix& P, const int* active, const float* x0, int n0, const float* x1, float* partial) {
  if (P.ntask <= 0) return;
  if (P.W != kW32 || P.per != kPer32 || !P.val32) {
    check(hipErrorInvalidValue, "panel_spmv: not an FP32 panel matrix", __FILE__, __LINE__);
    return;
  }
  launch<float, float, kW32, kPer32>(P, active, P.val32, x0, n0, x1, partial);
  LAUNCHCHK();
}

void panel_spmv(const PanelMatrix& P, const int* active, const double* x0, int n0, const double* x1,
                double* partial) {
  if (P.ntask <= 0) return;
  if (P.W != kW64 || P.per != kPer64 || !P.val64) {
    check(hipErrorInvalidValue, "panel_spmv: not an FP64 panel matrix", __FILE__, __LINE__);
    return;
  }
  launch<double, double, kW64, kPer64>(P, active, P.val64, x0, n0, x1, partial);
  LAUNCHCHK();
}

void panel_spmv_f32vals(const PanelMatrix& P, const float* val32, const int* active, const double* x0, int n0,
                        const double* x1, double* partial) {
  if (P.ntask <= 0) return;
  if (P.W != kW64 || P.per != kPer64 || !val32) {
    check(hipErrorInvalidValue, "panel_spmv_f32vals: not an FP64-built panel matrix", __FILE__, __LINE__);
    return;
  }
  launch<float, double, kW64, kPer64, 2>(P, active, val32, x0, n0, x1, partial);
  LAUNCHCHK();
}

float* panel_alloc_f32vals(const PanelMatrix& P) {
  const long long pad = 64LL * P.per;
  float* v = (float*)dev_alloc(sizeof(float) * (size_t)(P.nnz + pad));
  // the padding past nnz is read (and multiplied by zero): keep it finite
  if (v) HIPCHK(hipMemsetAsync(v + P.nnz, 0, sizeof(float) * (size_t)pad, rt().stream));
  return v;
}

void panel_fill_f32vals(const PanelMatrix& P, const double* csr_val, float* val32) {
  if (P.nnz <= 0) return;
  const int g = grid_for(P.nnz, kBlock, 16384);
  hipLaunchKernelGGL(k_panel_gather, dim3(g), dim3(kBlock), 0, rt().stream, P.nnz, (const int*)P.src, csr_val,
                     (double*)nullptr, val32);
  LAUNCHCHK();
}

int hybrid_build(HybridOp* H, int R, int C, int bits, const int* rp, const int* ci) {
  memset(H, 0, sizeof(*H));
  if (R <= 0 || C <= 0 || rp[R] <= 0) return 1;
  const long long total = rp[R];
  // Rp: start of the trailing run of rows with at most 4 entries
  int Rp = R;
  while (Rp > 0 && rp[Rp] - rp[Rp - 1] <= 4) --Rp;
  if (Rp == 0) return 1;
  const int W = panel_width(bits);
  const int nblk = (C + W - 1) / W;
  std::vector<long long> cnt(nblk, 0);
  for (long long k = 0; k < rp[Rp]; ++k) cnt[ci[k] / W]++;
  std::vector<unsigned char> sel(nblk, 0);
  long long pan_nnz = 0;
  int np = 0;
  for (int b = 0; b < nblk; ++b)
    if (cnt[b] >= Rp) {
      sel[b] = 1;
      pan_nnz += cnt[b];
      np++;
    }
  if (np == 0 || 2 * pan_nnz < total) return 1;
  const double vb = bits / 8.0;
  const double t_rb = total * (vb + 4.0) / 1.5;
  const double t_h = (pan_nnz * (vb + 2.0) + np * (Rp + 1.0) * 4.0 + 2.0 * np * (double)Rp * vb) / 5.0 +
                     (total - pan_nnz) * (vb + 8.0) / 1.5 + R * 8.0 / 5.0;
  if (t_h > 0.8 * t_rb) return 1;
  if (panel_build(&H->pan, Rp, C, bits, rp, ci, sel.data())) {
    panel_free(&H->pan);
    return -1;
  }
  // remainder CSR: entries of rows >= Rp and of unselected blocks, in order
  std::vector<int> rrp((size_t)R + 1, 0), rci, rsrc;
  rci.reserve((size_t)(total - pan_nnz));
  rsrc.reserve((size_t)(total - pan_nnz));
  for (int r = 0; r < R; ++r) {
    for (int k = rp[r]; k < rp[r + 1]; ++k)
      if (r >= Rp || !sel[ci[k] / W]) {
        rci.push_back(ci[k]);
        rsrc.push_back(k);
      }
    rrp[(size_t)r + 1] = (int)rci.size();
  }
  H->R = R;
  H->C = C;
  H->Rp = Rp;
  H->rnnz = (long long)rci.size();
  const size_t rn = (size_t)(H->rnnz > 0 ? H->rnnz : 1);
  H->rrp = (int*)dev_alloc(sizeof(int) * rrp.size());
  H->rci = (int*)dev_alloc(sizeof(int) * rn);
  H->rsrc = (int*)dev_alloc(sizeof(int) * rn);
  if (bits == 32) H->rv32 = (float*)dev_alloc(sizeof(float) * rn);
  else H->rv64 = (double*)dev_alloc(sizeof(double) * rn);
  if (!H->rrp || !H->rci || !H->rsrc || (bits == 32 ? !H->rv32 : !H->rv64)) {
    hybrid_free(H);
    return -1;
  }
  upload(H->rrp, rrp.data(), sizeof(int) * rrp.size());
  if (H->rnnz) {
    upload(H->rci, rci.data(), sizeof(int) * rci.size());
    upload(H->rsrc, rsrc.data(), sizeof(int) * rsrc.size());
  }
  H->ver = -1;
  return 0;
}

int hybrid_add_f32(HybridOp* H) {
  if (!H->pan.val64) return 1;
  if (!H->dc32 && !(H->dc32 = panel_alloc_f32vals(H->pan))) return 1;
  if (!H->rv32 && !(H->rv32 = (float*)dev_alloc(sizeof(float) * (size_t)(H->rnnz > 0 ? H->rnnz : 1)))) return 1;
  H->ver = -1;
  return 0;
}

void hybrid_free(HybridOp* H) {
  panel_free(&H->pan);
  dev_free(H->rrp);
  dev_free(H->rci);
  dev_free(H->rsrc);
  dev_free(H->rv32);
  dev_free(H->rv64);
  dev_free(H->dc32);
  memset(H, 0, sizeof(*H));
}

void hybrid_refresh_values(HybridOp* H, const double* csr_val) {
  panel_refresh_values(&H->pan, csr_val);
  if (H->dc32) panel_fill_f32vals(H->pan, csr_val, H->dc32);
  if (H->rnnz > 0) {
    const int g = grid_for(H->rnnz, kBlock, 16384);
    hipLaunchKernelGGL(k_panel_gather, dim3(g), dim3(kBlock), 0, rt().stream, H->rnnz, (const int*)H->rsrc, csr_val,
                       H->rv64, H->rv32);
    LAUNCHCHK();
  }
}

int panel_worth(long long nnz, int R, int C, int bits) {
  const int W = bits == 32 ? kW32 : kW64;
  const long long np = C > 0 ? (C + W - 1) / W : 1;
  if (R <= 0 || nnz <= 0) return 0;
  return np == 1 || nnz >= 4 * np * (long long)R;
}

double panel_spmv_bytes(const PanelMatrix& P, int value_bytes) {
  // x: the columns of the panels (every column when all blocks are panels)
  const double xcols = P.pcol ? (double)P.np * P.W : (double)P.C;
  return (double)P.nnz * (value_bytes + 2.0) + (double)P.np * (P.R + 1.0) * 4.0 + xcols * value_bytes +
         (double)P.np * P.R * value_bytes + (double)P.ntask * 24.0;
}

}  // namespace osqp_hip
