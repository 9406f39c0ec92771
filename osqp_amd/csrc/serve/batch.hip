// batch.hip — batched serving engine (include/osqp_amd_serve.h).
//
// Device side: one persistent kernel per size bucket and epoch.  Its
// workgroups take problems from the epoch's table with an atomic counter; a
// workgroup runs its problem's pending command (setup, or a new solve) and then
// up to `iters` ADMM iterations (team.h), writes the problem's state back to
// its HBM block and takes the next problem, until the table is exhausted.
//
// Host side: at osqp_batch_add a problem's stage image is built (CSR of A, A'
// and P with both triangles, row-group rows and offsets; in a pinned image
// arena when it has room), and the problem is admitted when a slot, an HBM
// block of the device arena and pinned staging memory are free; its setup
// epoch copies the image into its HBM block and derives the rest there.  Each size bucket runs its epochs
// on its own stream: when a bucket's epoch has ended the host collects the
// results the kernel wrote to pinned memory, writes the bucket's next table
// into pinned memory and replays the bucket's hipGraph (one kernel node; grid,
// team size and team LDS fixed per graph).  An online tuner tries team sizes and
// teams per CU on the first epochs of every bucket and keeps the fastest
// (ADMM iterations per second); results do not depend on that choice.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "osqp_amd_serve.h"
#include "team.h"

extern "C" OSQPInt osqp_algebra_init_libs(OSQPInt device);
extern "C" void osqp_algebra_free_libs(void);

namespace osqp_serve {

// ---- epoch kernel --------------------------------------------------------------
struct BEntry {
  BProb* dev;           // the problem's descriptor in HBM
  const BProb* init;    // host-mapped descriptor of a new problem (cmd 1)
  const double* upd_q;  // cmd 2 / 3: new q (raw) or null
  const BMatUpd* upd_mat;  // cmd 3: new matrix values (host-mapped)
  int cmd;              // 0 continue, 1 setup + solve, 2 new solve, 3 new matrix values + new solve
  int iters_done;       // written by the kernel
};

struct BTable {
  int count;  // entries
  int epoch;  // selects the work counter
  int iters;  // ADMM iterations per problem and epoch
  int pad;
  BEntry e[1];
};

constexpr int kCounterRing = 1024;

template <typename V, bool SL>
__device__ void run_entry(const Team& t, BProb& P, const BEntry& E, int iters, int* done_iters) {
  Solver<V, SL> sv(t, P);
  if (E.cmd == 1) {
    sv.setup();
  } else if (E.cmd == 2) {
    sv.new_solve(E.upd_q);
  } else if (E.cmd == 3) {
    sv.update_matrices(*E.upd_mat);
    sv.new_solve(E.upd_q);
  }
  int k = 0;
  if (sv.S.phase == kPhaseRun) {
    for (; k < iters;) {
      ++k;
      if (sv.iterate()) {
        sv.finish();
        break;
      }
    }
  }
  t.sync();
  if (t.tid == 0) {
    P.st = sv.S;
    *done_iters = k;
  }
}

// Teams of up to 512 threads.  The bound lets the compiler keep the solver's
// working set in registers (about 250 VGPRs, two waves per SIMD) instead of
// spilling it at 128 (a 1024-thread bound); on the MPC stream that cut the GPU
// time of the same work by 28 %.
// SL: with the FP32 CG's slices (size buckets kSlicedBucket and up)
template <typename V, bool SL>
__global__ __launch_bounds__(512) void k_epoch(BTable* tab, int* counters, int lds_bytes) {
  __shared__ double red[kMaxRedBlocks];
  __shared__ double red2[kMaxRedBlocks];
  extern __shared__ double dyn_lds[];
  __shared__ BProb P;
  __shared__ int slot, done_iters, decision;
  Team t;
  t.tid = threadIdx.x;
  t.T = blockDim.x;
  t.lane = threadIdx.x & (kWave - 1);
  t.wv = threadIdx.x / kWave;
  t.W = blockDim.x / kWave;
  t.red = red;
  t.red2 = red2;
  t.flag = &decision;
  t.lds = (char*)dyn_lds;
  t.lds_bytes = lds_bytes;
  const int ep = tab->epoch;
  const int count = tab->count;
  const int iters = tab->iters;
  int* ctr = counters + (ep % kCounterRing);
  // the next epoch's counter starts at zero (the next launch follows this one
  // on the same stream)
  if (blockIdx.x == 0 && threadIdx.x == 0) counters[(ep + 1) % kCounterRing] = 0;
  for (;;) {
    if (threadIdx.x == 0) slot = atomicAdd(ctr, 1);
    __syncthreads();
    const int s = slot;
    __syncthreads();
    if (s >= count) break;
    const BEntry E = tab->e[s];
    const BProb* src = E.cmd == 1 ? E.init : E.dev;
    const int* sw = (const int*)src;
    int* dw = (int*)&P;
    for (int i = threadIdx.x; i < (int)(sizeof(BProb) / 4); i += blockDim.x) dw[i] = sw[i];
    __syncthreads();
    run_entry<V, SL>(t, P, E, iters, &done_iters);
    __syncthreads();
    int* gw = (int*)E.dev;
    for (int i = threadIdx.x; i < (int)(sizeof(BProb) / 4); i += blockDim.x) gw[i] = dw[i];
    if (threadIdx.x == 0) tab->e[s].iters_done = done_iters;
    __syncthreads();
  }
}

// ---- host side ------------------------------------------------------------------
namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// copies bytes from host or device memory (callers may pass either)
void copy_in(void* dst, const void* src, size_t bytes) {
  if (osqp_amd_copy_to_host(dst, src, bytes)) fprintf(stderr, "osqp_batch: copy from device memory failed\n");
}

// A problem's arrays on the host: the caller's own where they are host
// memory; copies where any of them is device memory.
struct HostInputs {
  OSQPCscMatrix P{}, A{};
  const OSQPFloat *q = nullptr, *l = nullptr, *u = nullptr, *x0 = nullptr, *y0 = nullptr;
  std::vector<OSQPInt> pp, pi, ap, ai;
  std::vector<OSQPFloat> px, ax, vq, vl, vu, vx0, vy0;
};

// whether any of a problem's arrays is device memory (hipPointerGetAttributes
// serialises concurrent callers: call it from one thread)
bool any_device(const OSQPCscMatrix* P, const OSQPFloat* q, const OSQPCscMatrix* A, const OSQPFloat* l,
                const OSQPFloat* u, const OSQPFloat* x0, const OSQPFloat* y0) {
  const void* ptrs[] = {P->p, P->i, P->x, A->p, A->i, A->x, q, l, u, x0, y0};
  for (const void* v : ptrs)
    if (osqp_amd_ptr_is_device(v)) return true;
  return false;
}

// dev: any_device(...) of these arrays, found by the caller
bool host_inputs(const OSQPCscMatrix* P, const OSQPFloat* q, const OSQPCscMatrix* A, const OSQPFloat* l,
                 const OSQPFloat* u, const OSQPFloat* x0, const OSQPFloat* y0, bool dev, HostInputs* H) {
  H->P = *P;
  H->A = *A;
  H->q = q;
  H->l = l;
  H->u = u;
  H->x0 = x0;
  H->y0 = y0;
  if (!dev) return true;
  const size_t n = (size_t)P->n, m = (size_t)A->m;
  bool good = true;
  auto grab = [&](auto& vec, const auto* src, size_t count) {
    vec.resize(count ? count : 1);
    if (count) good = good && osqp_amd_copy_to_host(vec.data(), src, sizeof(vec[0]) * count) == 0;
    return vec.data();
  };
  H->P.p = grab(H->pp, P->p, n + 1);
  H->A.p = grab(H->ap, A->p, n + 1);
  if (!good) return false;
  H->P.i = grab(H->pi, P->i, (size_t)H->pp[n]);
  H->P.x = grab(H->px, P->x, (size_t)H->pp[n]);
  H->A.i = grab(H->ai, A->i, (size_t)H->ap[n]);
  H->A.x = grab(H->ax, A->x, (size_t)H->ap[n]);
  H->q = grab(H->vq, q, n);
  H->l = l ? grab(H->vl, l, m) : nullptr;
  H->u = u ? grab(H->vu, u, m) : nullptr;
  H->x0 = x0 ? grab(H->vx0, x0, n) : nullptr;
  H->y0 = y0 ? grab(H->vy0, y0, m) : nullptr;
  return good;
}

// first-fit allocator over [0, size) with coalescing
class RangeAlloc {
 public:
  void reset(size_t size) {
    free_.clear();
    if (size) free_[0] = size;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = align_up(bytes ? bytes : 1, 256);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes, start = it->first + bytes;
        free_.erase(it);
        if (rest) free_[start] = rest;
        used_[*off] = bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off) {
    auto u = used_.find(off);
    if (u == used_.end()) return;
    size_t start = off, len = u->second;
    used_.erase(u);
    auto next = free_.lower_bound(start);
    if (next != free_.end() && start + len == next->first) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        start = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    free_[start] = len;
  }

 private:
  std::map<size_t, size_t> free_, used_;
};

// Pinned memory for the stage images of problems not yet set up: the
// preparation (worker threads) builds a problem's image there, so admission
// does not copy it, and the block is released with the stage after the
// problem's setup epoch.  When it is full an image is built on the heap and
// copied into the stage at admission.  Separate from the staging pool, so
// images waiting in the queue never hold the pool space admission needs.
struct ImageArena {
  char* base = nullptr;
  size_t size = 0;
  RangeAlloc alloc;
  std::mutex mu;
  char* take(size_t bytes, size_t* off) {
    if (!base) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    return alloc.alloc(bytes, off) ? base + *off : nullptr;
  }
  void give(size_t off) {
    std::lock_guard<std::mutex> lk(mu);
    alloc.release(off);
  }
};

// Sizes of the row groups of one CSR matrix (the BCsr layout).  Rows of
// 1..kLongRow entries are packed in row order into groups of kGroupSlots
// entries; a row that does not fit the rest of the current group starts the
// next one (the rest is padding), so no row crosses a group.  With the FP32
// CG's slices (ell), the short rows (under kEllMin entries) are packed first,
// into groups [0, ng_short), and the medium rows (kEllMin..kLongRow) after
// them: the FP64 products walk every group, the FP32 CG only the short ones
// and takes the medium rows from the slices.
struct GroupDims {
  int ng = 0, nlong = 0, nempty = 0;
  long long long_slots = 0;  // long-row entries of the compact copy (rows padded to kLongAlign)
  int cols = 0;              // columns of the matrix
  bool ell = false;          // medium rows also in slices (FP32 engine, fewer than 65536 columns)
  int ng_short = 0;          // groups of the short rows (all groups without slices)
  int nslice = 0;            // slices of kWave medium rows
  long long ell_slots = 0;   // slice entries: kWave x the slice's width, summed
};

// Size bucket of a problem (one hipGraph and stream each): by
// w = n + m + nnz(A) + nnz(P), 4096 / 16384 / 65536 / 262144 / 2^20 / more.
inline int size_bucket(long long n, long long m, long long nnzA, long long nnzP) {
  const long long w = n + m + nnzA + nnzP;
  if (w <= (1 << 12)) return 0;
  if (w <= (1 << 14)) return 1;
  if (w <= (1 << 16)) return 2;
  if (w <= (1 << 18)) return 3;
  if (w <= (1 << 20)) return 4;
  return 5;
}
// Buckets from kSlicedBucket on run the kernel with the FP32 CG's slices,
// and only their problems get slices: a 5000-asset portfolio's CG iteration
// 93.5 -> 81 us with them, while a 2000-asset one's (bucket 2) went from
// 25 to 34 us, and MPC problems' epochs took 40 % longer in the kernel that
// holds the slice path at all.
constexpr int kSlicedBucket = 3;

// medium rows in slice order: longest first, row order among equals
void ell_order(const int* rp, int rows, std::vector<int>* order) {
  order->clear();
  for (int r = 0; r < rows; ++r) {
    const int len = rp[r + 1] - rp[r];
    if (len >= kEllMin && len <= kLongRow) order->push_back(r);
  }
  std::stable_sort(order->begin(), order->end(),
                   [&](int a, int b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; });
}

// a slice's width: its longest row rounded up to whole kEllVec chunks
inline int ell_width(int len) { return (len + kEllVec - 1) / kEllVec * kEllVec; }

GroupDims count_groups(const int* rp, int rows, int cols, bool fp32) {
  GroupDims d;
  d.cols = cols;
  int nmed = 0;
  for (int r = 0; r < rows; ++r) {
    const int len = rp[r + 1] - rp[r];
    if (len >= kEllMin && len <= kLongRow) nmed++;
  }
  d.ell = fp32 && cols < 65536 && nmed >= kEllMinRows;
  int fill = 0;
  auto pack = [&](int len) {
    if (d.ng == 0 || fill + len > kGroupSlots) {
      d.ng++;
      fill = 0;
    }
    fill += len;
  };
  for (int r = 0; r < rows; ++r) {
    const int len = rp[r + 1] - rp[r];
    if (len > kLongRow) {
      d.nlong++;
      d.long_slots += (len + kLongAlign - 1) / kLongAlign * kLongAlign;
    } else if (len == 0) {
      d.nempty++;
    } else if (!d.ell || len < kEllMin) {
      pack(len);
    }
  }
  d.ng_short = d.ng;
  if (d.ell) {
    // the medium rows' groups start a new group
    fill = kGroupSlots;
    for (int r = 0; r < rows; ++r) {
      const int len = rp[r + 1] - rp[r];
      if (len >= kEllMin && len <= kLongRow) pack(len);
    }
    thread_local std::vector<int> order;
    ell_order(rp, rows, &order);
    d.nslice = (nmed + kWave - 1) / kWave;
    for (int s = 0; s < d.nslice; ++s) {
      const int r = order[(size_t)s * kWave];
      d.ell_slots += (long long)kWave * ell_width(rp[r + 1] - rp[r]);
    }
  }
  return d;
}

// Dimensions of a problem and its stage image, built once at add time in its
// final byte layout, in pinned memory when the image arena has room (then
// admission copies nothing), else on the heap (then one copy into the stage).
struct HostData {
  int n = 0, m = 0;
  int nnzP = 0;  // the full symmetric P
  int nnzA = 0;
  GroupDims gP, gA, gT;
  bool warm_x = false, warm_y = false;
  std::unique_ptr<char[]> img;  // stage bytes [st_static, st_total): static image, x0, y0
  bool img_pinned = false;      // the image is at img_off in the engine's image arena instead
  size_t img_off = 0;
  char* img_ptr = nullptr;      // the image's first byte, wherever it lives
  // resident problems: where each entry of the input CSC of P (triangle) and
  // of A lives in the HBM copies, for value updates (osqp_update_data_mat):
  // P entry k at row-of-column slot pslot_lo[k] and, off the diagonal, at
  // the mirrored slot pslot_up[k] (-1 on the diagonal); A entry k at CSR slot
  // aslot[k] (and at slot k of A', which is the CSC itself)
  std::vector<int> pslot_lo, pslot_up, aslot;
  int nnzP_in = 0;
};

// byte layout of a problem's HBM block and of its pinned images
struct Layout {
  size_t desc;  // BProb at offset 0
  // static image (copied from the pinned stage at setup): offsets from block start
  size_t s_begin, s_end;
  size_t Prp, Pci, Pval, Pdiag, Arp, Aci, Aval, Trp, Tci, Tval, q, l, u;
  struct Groups {
    // static (in the stage image): longs, empty, goff, grow, loff, eoff, erl;
    // dynamic (derived on the GPU at setup): gci, gval, lci16, lv32, eci16, ev32
    size_t longs, empty, goff, grow, gci, gval;
    // compact copy of the long rows for the FP32 CG (offsets 0 when absent)
    bool compact;
    size_t loff, lci16, lv32;
    // the FP32 CG's slices of the medium rows (offsets 0 when absent)
    size_t eoff, erl, eci16, ev32;
  } gP, gA, gT;
  // dynamic
  size_t ctype, D, Dinv, E, Einv, rho, rhoinv, xa, xb, za, zb, y, xt, zt, dx, dy, Ax, Px, Aty, tn, tm, tn2;
  size_t x64, r64, rhs64, w64, ax64, prec, cd, cp, cKp, cr, cy, cw, cdinv;
  size_t total;
  // pinned: init descriptor + static image + warm start; results (+ q buffer)
  size_t st_total, st_static, st_x0, st_y0;
  size_t rs_total, rs_x, rs_y, rs_pc, rs_dc, rs_res, rs_q;
};

Layout make_layout(const HostData& h, int vbytes) {
  Layout L{};
  const size_t n = h.n, m = h.m;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    o = align_up(o, 64);
    const size_t at = o;
    o += bytes ? bytes : 8;
    return at;
  };
  L.desc = take(sizeof(BProb));
  o = align_up(o, 64);
  L.s_begin = o;
  L.Prp = take(4 * (n + 1));
  L.Pci = take(4 * (size_t)h.nnzP);
  L.Pval = take(8 * (size_t)h.nnzP);
  L.Pdiag = take(4 * n);
  L.Arp = take(4 * (m + 1));
  L.Aci = take(4 * (size_t)h.nnzA);
  L.Aval = take(8 * (size_t)h.nnzA);
  L.Trp = take(4 * (n + 1));
  L.Tci = take(4 * (size_t)h.nnzA);
  L.Tval = take(8 * (size_t)h.nnzA);
  auto groups = [&](Layout::Groups* o, const GroupDims& g) {
    const size_t slots = (size_t)g.ng * kGroupSlots;
    o->longs = take(4 * (size_t)g.nlong);
    o->empty = take(4 * (size_t)g.nempty);
    o->goff = take(slots);
    o->grow = take(4 * slots);
    // the compact long rows: FP32 engine, columns in 16 bits
    o->compact = vbytes == 4 && g.cols < 65536 && g.nlong > 0;
    if (o->compact) o->loff = take(4 * ((size_t)g.nlong + 1));
    if (g.ell) {
      o->eoff = take(4 * ((size_t)g.nslice + 1));
      o->erl = take(8 * (size_t)g.nslice * kWave);
    }
  };
  groups(&L.gP, h.gP);
  groups(&L.gA, h.gA);
  groups(&L.gT, h.gT);
  L.q = take(8 * n);
  L.l = take(8 * m);
  L.u = take(8 * m);
  o = align_up(o, 64);
  L.s_end = o;
  auto dyn = [&](Layout::Groups* o, const GroupDims& g) {
    o->gci = take(4 * (size_t)g.ng * kGroupSlots);
    o->gval = take(8 * (size_t)g.ng * kGroupSlots);
    if (o->compact) {
      o->lci16 = take(2 * (size_t)g.long_slots);
      o->lv32 = take(4 * (size_t)g.long_slots);
    }
    if (g.ell) {
      o->eci16 = take(2 * (size_t)g.ell_slots);
      o->ev32 = take(4 * (size_t)g.ell_slots);
    }
  };
  dyn(&L.gP, h.gP);
  dyn(&L.gA, h.gA);
  dyn(&L.gT, h.gT);
  L.ctype = take(4 * m);
  L.D = take(8 * n);
  L.Dinv = take(8 * n);
  L.E = take(8 * m);
  L.Einv = take(8 * m);
  L.rho = take(8 * m);
  L.rhoinv = take(8 * m);
  L.xa = take(8 * n);
  L.xb = take(8 * n);
  L.za = take(8 * m);
  L.zb = take(8 * m);
  L.y = take(8 * m);
  L.xt = take(8 * n);
  L.zt = take(8 * m);
  L.dx = take(8 * n);
  L.dy = take(8 * m);
  L.Ax = take(8 * m);
  L.Px = take(8 * n);
  L.Aty = take(8 * n);
  L.tn = take(8 * n);
  L.tm = take(8 * m);
  L.tn2 = take(8 * n);
  L.x64 = take(8 * n);
  L.r64 = take(8 * n);
  L.rhs64 = take(8 * n);
  L.w64 = take(8 * m);
  L.ax64 = take(8 * m);
  L.prec = take(8 * n);
  L.cd = take(vbytes * n);
  L.cp = take(vbytes * n);
  L.cKp = take(vbytes * n);
  L.cr = take(vbytes * n);
  L.cy = take(vbytes * n);
  L.cw = take(vbytes * m);
  L.cdinv = take(vbytes * n);
  L.total = align_up(o, 256);
  // pinned stage: [init descriptor][static image][x0][y0]
  size_t s = align_up(sizeof(BProb), 64);
  L.st_static = s;
  s += L.s_end - L.s_begin;
  s = align_up(s, 64);
  L.st_x0 = s;
  s += h.warm_x ? 8 * n : 0;
  s = align_up(s, 64);
  L.st_y0 = s;
  s += h.warm_y ? 8 * m : 0;
  L.st_total = align_up(s, 256);
  // pinned results: [x][y][prim_inf_cert][dual_inf_cert][BResult][q buffer]
  size_t r = 0;
  L.rs_x = r;
  r = align_up(r + 8 * (n ? n : 1), 64);
  L.rs_y = r;
  r = align_up(r + 8 * (m ? m : 1), 64);
  L.rs_pc = r;
  r = align_up(r + 8 * (m ? m : 1), 64);
  L.rs_dc = r;
  r = align_up(r + 8 * (n ? n : 1), 64);
  L.rs_res = r;
  r = align_up(r + sizeof(BResult), 64);
  L.rs_q = r;
  r = align_up(r + 8 * (n ? n : 1), 64);
  L.rs_total = align_up(r, 256);
  return L;
}

// Pass 2 of the packing: the long and empty row lists, each group slot's row
// and its entry's offset within the row (a group's entry number e sits in
// lane e / kGroupPer, element e % kGroupPer, at slot e: a lane's slots are
// contiguous), the long rows' offsets in the compact copy and the slices'
// offsets and (row, length) pairs, written into the image img whose first
// byte is block offset base.  Columns and values are derived from these and
// the CSR arrays on the GPU at setup (Solver::fill_group_values).
void write_groups(const int* rp, int rows, const GroupDims& d, const Layout::Groups& o, char* img, size_t base) {
  int* longs = (int*)(img + (o.longs - base));
  int* empty = (int*)(img + (o.empty - base));
  unsigned char* go0 = (unsigned char*)(img + (o.goff - base));
  int* gr0 = (int*)(img + (o.grow - base));
  const size_t slots = (size_t)d.ng * kGroupSlots;
  std::fill(go0, go0 + slots, (unsigned char)0);
  std::fill(gr0, gr0 + slots, -1);
  int* loff = o.compact ? (int*)(img + (o.loff - base)) : nullptr;
  int nl = 0, ne = 0, grp = -1, fill = 0, lpos = 0;
  auto pack = [&](int r) {
    const int len = rp[r + 1] - rp[r];
    if (grp < 0 || fill + len > kGroupSlots) {
      grp++;
      fill = 0;
    }
    unsigned char* go = go0 + (size_t)grp * kGroupSlots;
    int* gr = gr0 + (size_t)grp * kGroupSlots;
    for (int k = 0; k < len; ++k) {
      const int slot = fill++;
      go[slot] = (unsigned char)k;
      gr[slot] = r;
    }
  };
  for (int r = 0; r < rows; ++r) {
    const int len = rp[r + 1] - rp[r];
    if (len > kLongRow) {
      if (o.compact) {
        loff[nl] = lpos;
        lpos += (len + kLongAlign - 1) / kLongAlign * kLongAlign;
      }
      longs[nl++] = r;
      continue;
    }
    if (len == 0) {
      empty[ne++] = r;
      continue;
    }
    if (!d.ell || len < kEllMin) pack(r);
  }
  if (o.compact) loff[nl] = lpos;
  if (!d.ell) return;
  // the medium rows: their groups (a new group first), then the slices.
  // Slice s holds rows order[kWave s + lane] (row -1, length 0 past the
  // last), as (row, length) pairs; its width K is its first (longest) row's
  // rounded up to kEllVec; entry k of the lane's row sits at eoff[s] +
  // (k / kEllVec) kWave kEllVec + lane kEllVec + k % kEllVec, so a lane
  // loads kEllVec consecutive entries at a time.
  fill = kGroupSlots;
  for (int r = 0; r < rows; ++r) {
    const int len = rp[r + 1] - rp[r];
    if (len >= kEllMin && len <= kLongRow) pack(r);
  }
  int* eoff = (int*)(img + (o.eoff - base));
  int* erl = (int*)(img + (o.erl - base));
  thread_local std::vector<int> order;
  ell_order(rp, rows, &order);
  long long pos = 0;
  for (int sl = 0; sl < d.nslice; ++sl) {
    eoff[sl] = (int)pos;
    const int r0 = order[(size_t)sl * kWave];
    const int K = ell_width(rp[r0 + 1] - rp[r0]);
    for (int lane = 0; lane < kWave; ++lane) {
      const size_t at = (size_t)sl * kWave + lane;
      const int r = at < order.size() ? order[at] : -1;
      erl[2 * at] = r;
      erl[2 * at + 1] = r >= 0 ? rp[r + 1] - rp[r] : 0;
    }
    pos += (long long)kWave * K;
  }
  eoff[d.nslice] = (int)pos;
}

// Builds a problem's stage image (values raw, unscaled): A (CSC, m x n) ->
// A CSR and A' CSR (= the CSC arrays); P (CSC upper triangle) -> the full
// symmetric CSR with sorted columns and diagonal positions; bounds clamped to
// +-OSQP_INFTY as osqp_setup does; the row groups of all three; the warm start.
Layout build_host(HostData* h, int vbytes, const OSQPCscMatrix* P, const OSQPFloat* q, const OSQPCscMatrix* A,
                  const OSQPFloat* l, const OSQPFloat* u, const OSQPFloat* x0, const OSQPFloat* y0,
                  bool keep_maps, ImageArena* ia = nullptr) {
  const int n = P->n, m = A->m;
  h->n = n;
  h->m = m;
  h->nnzA = A->p[n];
  h->nnzP_in = P->p[n];
  if (keep_maps) {
    h->pslot_lo.assign((size_t)h->nnzP_in, -1);
    h->pslot_up.assign((size_t)h->nnzP_in, -1);
    h->aslot.assign((size_t)h->nnzA, -1);
  }
  // row pointers first: their sizes fix the layout (per-thread scratch keeps
  // a stream of adds free of allocations)
  thread_local std::vector<int> arp, prp, next;
  arp.assign((size_t)m + 1, 0);
  for (int k = 0; k < h->nnzA; ++k) arp[(size_t)A->i[k] + 1]++;
  for (int r = 0; r < m; ++r) arp[(size_t)r + 1] += arp[r];
  prp.assign((size_t)n + 1, 0);
  for (int c = 0; c < n; ++c)
    for (int k = P->p[c]; k < P->p[c + 1]; ++k) {
      const int r = P->i[k];
      prp[(size_t)c + 1]++;
      if (r != c) prp[(size_t)r + 1]++;
    }
  for (int r = 0; r < n; ++r) prp[(size_t)r + 1] += prp[r];
  h->nnzP = prp[n];
  const bool sliced = vbytes == 4 && size_bucket(n, m, h->nnzA, h->nnzP) >= kSlicedBucket;
  h->gP = count_groups(prp.data(), n, n, sliced);
  h->gA = count_groups(arp.data(), m, n, sliced);
  h->gT = count_groups(A->p, n, m, sliced);
  h->warm_x = x0 != nullptr;
  h->warm_y = y0 != nullptr;
  const Layout L = make_layout(*h, vbytes);
  // the image in the pinned arena when it has room, else on the heap
  char* img = ia ? ia->take(L.st_total - L.st_static, &h->img_off) : nullptr;
  h->img_pinned = img != nullptr;
  if (!img) {
    h->img.reset(new char[L.st_total - L.st_static]);
    img = h->img.get();
  }
  h->img_ptr = img;
  const size_t base = L.s_begin;
  auto at = [&](size_t off) { return img + (off - base); };
  const size_t nnzA = (size_t)h->nnzA, nnzP = (size_t)h->nnzP;
  // A' = the CSC arrays
  memcpy(at(L.Trp), A->p, 4 * ((size_t)n + 1));
  memcpy(at(L.Tci), A->i, 4 * nnzA);
  memcpy(at(L.Tval), A->x, 8 * nnzA);
  // A CSR by a counting sort on rows (columns ascending)
  int* Arp = (int*)at(L.Arp);
  int* Aci = (int*)at(L.Aci);
  double* Aval = (double*)at(L.Aval);
  memcpy(Arp, arp.data(), 4 * ((size_t)m + 1));
  next.assign(arp.begin(), arp.end() - 1);
  for (int j = 0; j < n; ++j)
    for (int k = A->p[j]; k < A->p[j + 1]; ++k) {
      const int pos = next[A->i[k]]++;
      Aci[pos] = j;
      Aval[pos] = A->x[k];
      if (keep_maps) h->aslot[(size_t)k] = pos;
    }
  // P full: row r = column r of the triangle (rows i <= r) then the entries
  // (r, c) of the triangle's columns c > r, ascending
  int* Prp = (int*)at(L.Prp);
  int* Pci = (int*)at(L.Pci);
  double* Pval = (double*)at(L.Pval);
  int* Pdiag = (int*)at(L.Pdiag);
  memcpy(Prp, prp.data(), 4 * ((size_t)n + 1));
  std::fill(Pdiag, Pdiag + n, -1);
  next.assign(prp.begin(), prp.end() - 1);
  for (int c = 0; c < n; ++c)  // lower part of row c: column c of the triangle (entries (i, c), i <= c)
    for (int k = P->p[c]; k < P->p[c + 1]; ++k) {
      const int pos = next[c]++;
      Pci[pos] = P->i[k];
      Pval[pos] = P->x[k];
      if (P->i[k] == c) Pdiag[c] = pos;
      if (keep_maps) h->pslot_lo[(size_t)k] = pos;
    }
  for (int c = 0; c < n; ++c)  // upper part of row r: (r, c) for c > r
    for (int k = P->p[c]; k < P->p[c + 1]; ++k) {
      const int r = P->i[k];
      if (r == c) continue;
      const int pos = next[r]++;
      Pci[pos] = c;
      Pval[pos] = P->x[k];
      if (keep_maps) h->pslot_up[(size_t)k] = pos;
    }
  // a triangle column's rows are not required to be sorted: sort each row
  for (int r = 0; r < n; ++r) {
    const int s = Prp[r], e = Prp[r + 1];
    bool sorted = true;
    for (int k = s + 1; k < e && sorted; ++k) sorted = Pci[k - 1] < Pci[k];
    if (!sorted) {
      std::vector<std::pair<int, int>> row;  // (column, old slot)
      std::vector<double> v(Pval + s, Pval + e);
      for (int k = s; k < e; ++k) row.emplace_back(Pci[k], k);
      std::sort(row.begin(), row.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
        return a.first < b.first;
      });
      std::vector<int> moved_to((size_t)(e - s));
      for (int k = s; k < e; ++k) {
        Pci[k] = row[k - s].first;
        Pval[k] = v[(size_t)(row[k - s].second - s)];
        moved_to[(size_t)(row[k - s].second - s)] = k;
        if (Pci[k] == r) Pdiag[r] = k;
      }
      if (keep_maps)
        for (int k = 0; k < h->nnzP_in; ++k) {
          int& lo = h->pslot_lo[(size_t)k];
          int& hi = h->pslot_up[(size_t)k];
          if (lo >= s && lo < e) lo = moved_to[(size_t)(lo - s)];
          if (hi >= s && hi < e) hi = moved_to[(size_t)(hi - s)];
        }
    }
  }
  (void)nnzP;
  memcpy(at(L.q), q, 8 * (size_t)n);
  double* lo = (double*)at(L.l);
  double* up = (double*)at(L.u);
  for (int i = 0; i < m; ++i) {
    lo[i] = l[i] < -kInfty ? -kInfty : (l[i] > kInfty ? kInfty : l[i]);
    up[i] = u[i] < -kInfty ? -kInfty : (u[i] > kInfty ? kInfty : u[i]);
  }
  write_groups(Prp, n, h->gP, L.gP, img, base);
  write_groups(Arp, m, h->gA, L.gA, img, base);
  write_groups(A->p, n, h->gT, L.gT, img, base);
  if (x0) memcpy(img + (L.st_x0 - L.st_static), x0, 8 * (size_t)n);
  if (y0) memcpy(img + (L.st_y0 - L.st_static), y0, 8 * (size_t)m);
  return L;
}

enum : int { kQueued = 0, kActive = 1, kEnded = 2, kIdle = 3 };

struct HProb {
  int id = 0, bucket = 0, state = kQueued, cmd = 1, resident = 0;
  int qnew = 0;  // a resolve brought a new q
  HostData h;
  Layout L{};
  bool has_dev = false, has_stage = false, has_res = false;
  size_t dev_off = 0, stage_off = 0, res_off = 0;
  double t_req = 0.0, latency = 0.0;
  long long iters_since_add = 0;
  // matrix values staged by osqp_batch_update_data_mat for the next resolve
  std::vector<int> up_pslot, up_aslot, up_tslot;
  std::vector<double> up_pval, up_aval;
  bool has_upd = false;  // pinned record of the running cmd 3
  size_t upd_off = 0;
};

constexpr int kBuckets = 6;
// ADMM iterations per epoch: long epochs while a bucket holds more problems
// than twice the CU count (the launch cost is spread over more work), short
// ones below that, so results come back sooner.  A problem's iterates do not
// depend on where epochs end.  On the MPC stream (4096 requests), short epochs
// throughout cost 3 % of the saturated throughput and cut the latency at 70 %
// load from p50 64 / p99 95 ms to p50 42 / p99 62 ms.
constexpr int kEpochIters = 25;
constexpr int kEpochItersLight = 15;

int bucket_of(const HostData& h) { return size_bucket(h.n, h.m, h.nnzA, h.nnzP); }

struct Cand {
  int T, tpc;
};
const Cand kCands[] = {{64, 16}, {128, 8}, {256, 4}, {256, 8}, {512, 2}, {512, 4}};
constexpr int kNumCands = sizeof(kCands) / sizeof(kCands[0]);
// The tuner times every candidate of a bucket on two epochs, sweeping the
// list forward and then backward, and ranks them by the sum of their two
// rates: a bucket's load drifts while it is tuned (a burst drains, a stream
// ramps up), and a drift that is linear in time then weighs every candidate
// alike.  Buckets 0 and 1 (n + m + nnz(A) + nnz(P) <= 16384: at most 32 row
// groups a product) leave out the 512-thread teams, whose eight wavefronts
// such problems cannot keep busy; their epochs' times also move with the
// other buckets' epochs running beside them, which can rank such a team first.
constexpr int kSmallBuckets = 2;
inline int cands_of(int bucket) { return bucket < kSmallBuckets ? 4 : kNumCands; }
inline int trials_of(int bucket) { return 2 * cands_of(bucket); }
inline int cand_of_trial(int bucket, int t) {
  const int nc = cands_of(bucket);
  return t < nc ? t : 2 * nc - 1 - t;
}

struct Bucket {
  std::vector<int> active;
  BTable* tab = nullptr;  // pinned
  size_t tab_cap = 0;
  int* counters = nullptr;  // device
  int epoch = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int g_T = 0, g_tpc = 0;
  hipStream_t stream = nullptr;  // buckets run side by side, each on its own stream
  bool inflight = false;         // an epoch of this bucket is on the GPU
  std::vector<int> incoming;     // admitted or re-solved while it was in flight
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // tuner
  int trial = 0;  // next trial (cand_of_trial gives its candidate); trials_of(bucket) when tuned
  double rate[kNumCands] = {0};
  int best = 2;   // 256 x 4 until tuned
  int fixT = 0, fixTpc = 0;  // pinned configuration (osqp_batch_set_team)
  bool sliced = false;       // runs the kernel with slices (bucket >= kSlicedBucket)
};

}  // namespace

struct Engine {
  BSettings set{};
  int bits = 32;
  int max_active = 4096;
  int cus = 256;
  size_t lds_per_cu = 65536;
  // per kernel (FP32, FP32 with slices, FP64): team size and dynamic LDS
  // bytes of the last lookup, and the kernel's dynamic-LDS limit as set (only
  // ever raised: the graphs of other buckets may launch with the larger size)
  int lds_T[3] = {0, 0, 0};
  int lds_B[3] = {0, 0, 0};
  int lds_max[3] = {0, 0, 0};
  char* arena = nullptr;
  size_t arena_size = 0;
  RangeAlloc arena_alloc;
  char* pool = nullptr;  // pinned
  size_t pool_size = 0;
  RangeAlloc pool_alloc;
  ImageArena images;  // pinned; the stage images of problems not yet set up
  std::map<int, HProb*> probs;
  std::vector<int> queue;  // waiting for admission, in arrival order
  std::vector<int> last_done;  // solves that ended in the last epoch
  // images of the problems admitted in this step, copied into their stages
  // by flush_stage_copies (on several threads when large)
  struct StageCopy {
    char* dst;
    std::unique_ptr<char[]> src;
    size_t bytes;
  };
  std::vector<StageCopy> stage_copies;
  int next_id = 1;
  int n_active = 0;
  Bucket buckets[kBuckets];
  long long epochs = 0, admm_iters = 0;
  double gpu_s = 0.0;
  bool failed = false;
  std::vector<int> resumed;       // resident problems re-solved since the last step
  hipEvent_t ref = nullptr;       // time origin of the busy-time account
  double busy_until_ms = 0.0;     // end of the GPU busy intervals counted so far
};

namespace {

bool ok(Engine* E, hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  fprintf(stderr, "osqp_batch: %s: %s\n", what, hipGetErrorString(e));
  E->failed = true;
  return false;
}

void free_problem_memory(Engine* E, HProb* p, bool dev, bool stage, bool res) {
  if (dev && p->has_dev) {
    E->arena_alloc.release(p->dev_off);
    p->has_dev = false;
  }
  if (stage && p->has_stage) {
    E->pool_alloc.release(p->stage_off);
    p->has_stage = false;
  }
  if (stage && p->h.img_pinned) {
    E->images.give(p->h.img_off);
    p->h.img_pinned = false;
    p->h.img_ptr = nullptr;
  }
  if (res && p->has_res) {
    E->pool_alloc.release(p->res_off);
    p->has_res = false;
  }
}

// reserve memory and write the pinned images of a queued problem
bool admit(Engine* E, HProb* p) {
  HostData& h = p->h;
  const Layout& L = p->L;
  size_t d, s, r;
  if (!E->arena_alloc.alloc(L.total, &d)) return false;
  // the stage: descriptor, then the image unless it is already pinned
  if (!E->pool_alloc.alloc(h.img_pinned ? L.st_static : L.st_total, &s)) {
    E->arena_alloc.release(d);
    return false;
  }
  if (!E->pool_alloc.alloc(L.rs_total, &r)) {
    E->arena_alloc.release(d);
    E->pool_alloc.release(s);
    return false;
  }
  p->dev_off = d;
  p->stage_off = s;
  p->res_off = r;
  p->has_dev = p->has_stage = p->has_res = true;
  char* dev = E->arena + d;
  char* st = E->pool + s;
  char* rs = E->pool + r;
  // static image (laid out as in the HBM block) and warm start, built at add:
  // read where it is when pinned, else copied into the stage by
  // flush_stage_copies before the epoch is launched
  char* img = h.img_pinned ? h.img_ptr : st + L.st_static;
  if (!h.img_pinned) E->stage_copies.push_back({img, std::move(h.img), L.st_total - L.st_static});
  // descriptor
  BProb* P = (BProb*)st;
  memset(P, 0, sizeof(BProb));
  P->n = h.n;
  P->m = h.m;
  P->set = E->set;
  P->st.phase = kPhaseSetup;
  P->st.rho = E->set.rho0;
  auto csr = [&](BCsr* M, int rows, const GroupDims& g, const Layout::Groups& o, size_t rp, size_t ci,
                 size_t val) {
    M->rows = rows;
    M->ng = g.ng;
    M->nlong = g.nlong;
    M->nempty = g.nempty;
    M->rp = (int*)(dev + rp);
    M->ci = (int*)(dev + ci);
    M->val = (double*)(dev + val);
    M->longs = (int*)(dev + o.longs);
    M->empty = (int*)(dev + o.empty);
    M->goff = (unsigned char*)(dev + o.goff);
    M->grow = (int*)(dev + o.grow);
    M->gci = (int*)(dev + o.gci);
    M->gval = (double*)(dev + o.gval);
    M->loff = o.compact ? (int*)(dev + o.loff) : nullptr;
    M->lci16 = o.compact ? (unsigned short*)(dev + o.lci16) : nullptr;
    M->lv32 = o.compact ? (float*)(dev + o.lv32) : nullptr;
    M->ng32 = g.ell ? g.ng_short : g.ng;
    M->nslice = g.ell ? g.nslice : 0;
    M->eoff = g.ell ? (int*)(dev + o.eoff) : nullptr;
    M->erl = g.ell ? (int*)(dev + o.erl) : nullptr;
    M->eci16 = g.ell ? (unsigned short*)(dev + o.eci16) : nullptr;
    M->ev32 = g.ell ? (float*)(dev + o.ev32) : nullptr;
  };
  csr(&P->P, h.n, h.gP, L.gP, L.Prp, L.Pci, L.Pval);
  csr(&P->A, h.m, h.gA, L.gA, L.Arp, L.Aci, L.Aval);
  csr(&P->At, h.n, h.gT, L.gT, L.Trp, L.Tci, L.Tval);
  P->Pdiag = (int*)(dev + L.Pdiag);
  P->ctype = (int*)(dev + L.ctype);
#define DPTR(f) P->f = (double*)(dev + L.f)
  DPTR(q);
  DPTR(l);
  DPTR(u);
  DPTR(D);
  DPTR(Dinv);
  DPTR(E);
  DPTR(Einv);
  DPTR(rho);
  DPTR(rhoinv);
  DPTR(xa);
  DPTR(xb);
  DPTR(za);
  DPTR(zb);
  DPTR(y);
  DPTR(xt);
  DPTR(zt);
  DPTR(dx);
  DPTR(dy);
  DPTR(Ax);
  DPTR(Px);
  DPTR(Aty);
  DPTR(tn);
  DPTR(tm);
  DPTR(tn2);
  DPTR(x64);
  DPTR(r64);
  DPTR(rhs64);
  DPTR(w64);
  DPTR(ax64);
  DPTR(prec);
#undef DPTR
  P->cd = dev + L.cd;
  P->cp = dev + L.cp;
  P->cKp = dev + L.cKp;
  P->cr = dev + L.cr;
  P->cy = dev + L.cy;
  P->cw = dev + L.cw;
  P->cdinv = dev + L.cdinv;
  P->stage = img;
  P->stage_bytes = (long long)(L.s_end - L.s_begin);
  P->arena_static = dev + L.s_begin;
  P->upd_q = nullptr;
  P->warm_x = h.warm_x ? (const double*)(img + (L.st_x0 - L.st_static)) : nullptr;
  P->warm_y = h.warm_y ? (const double*)(img + (L.st_y0 - L.st_static)) : nullptr;
  P->out_x = (double*)(rs + L.rs_x);
  P->out_y = (double*)(rs + L.rs_y);
  P->out_pcert = (double*)(rs + L.rs_pc);
  P->out_dcert = (double*)(rs + L.rs_dc);
  // certificates start at zero, as osqp_setup's calloc'd solution does (a
  // non-convex stop leaves them as they are)
  memset(P->out_pcert, 0, sizeof(double) * (size_t)h.m);
  memset(P->out_dcert, 0, sizeof(double) * (size_t)h.n);
  P->out = (BResult*)(rs + L.rs_res);
  memset(P->out, 0, sizeof(BResult));
  p->cmd = 1;
  return true;
}

// The stage copies of this step's admissions (hundreds of KB a problem, a
// few hundred problems a step: split over threads when over 2 MB).
void flush_stage_copies(Engine* E) {
  std::vector<Engine::StageCopy>& cs = E->stage_copies;
  if (cs.empty()) return;
  size_t total = 0;
  for (const auto& c : cs) total += c.bytes;
  const unsigned hw = std::thread::hardware_concurrency();
  const int nthreads = total < ((size_t)2 << 20) ? 1 : (int)std::min<size_t>(std::min<size_t>(8, hw ? hw : 1), cs.size());
  auto work = [&](int t) {
    for (size_t k = (size_t)t; k < cs.size(); k += (size_t)nthreads) memcpy(cs[k].dst, cs[k].src.get(), cs[k].bytes);
  };
  std::vector<std::thread> pool;
  try {
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(work, t);
  } catch (...) {
    // no more threads: this one copies what the missing threads would have
    for (int t = 1 + (int)pool.size(); t < nthreads; ++t) work(t);
  }
  work(0);
  for (auto& th : pool) th.join();
  cs.clear();
}

bool ensure_table(Engine* E, Bucket& B, size_t count) {
  if (B.tab && B.tab_cap >= count) return true;
  if (B.tab) {
    if (!ok(E, hipHostFree(B.tab), "hipHostFree")) return false;
    B.tab = nullptr;
  }
  size_t cap = std::max<size_t>(count, 256);
  cap = std::max<size_t>(cap, (size_t)E->max_active);
  if (!ok(E, hipHostMalloc((void**)&B.tab, sizeof(BTable) + sizeof(BEntry) * cap, hipHostMallocDefault),
          "hipHostMalloc table"))
    return false;
  B.tab_cap = cap;
  // the graph holds the table address
  if (B.exec) {
    hipGraphExecDestroy(B.exec);
    B.exec = nullptr;
  }
  if (B.graph) {
    hipGraphDestroy(B.graph);
    B.graph = nullptr;
  }
  return true;
}

// Dynamic LDS of a team of T threads: the CU's LDS shared by the teams of
// that size its registers let it hold (static LDS of each taken off), at least
// kTeamLdsBytes and at most what keeps that team count resident.  A problem
// whose CG vectors fit keeps them there (same arithmetic either way: the
// bits do not depend on it).  -1 on a HIP error.
const void* epoch_fn(const Engine* E, bool sliced) {
  if (E->bits != 32) return (const void*)k_epoch<double, false>;
  return sliced ? (const void*)k_epoch<float, true> : (const void*)k_epoch<float, false>;
}

int team_lds_bytes(Engine* E, int T, bool sliced) {
  const int w = E->bits == 32 ? (sliced ? 1 : 0) : 2;
  if (E->lds_T[w] == T) return E->lds_B[w];
  const void* fn = epoch_fn(E, sliced);
  hipFuncAttributes fa;
  int per_cu = 0;
  if (!ok(E, hipFuncGetAttributes(&fa, fn), "hipFuncGetAttributes") ||
      !ok(E, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, T, kTeamLdsBytes), "occupancy"))
    return -1;
  int bytes = kTeamLdsBytes;
  if (per_cu > 0) {
    const long long share = (long long)(E->lds_per_cu / (size_t)per_cu) - (long long)fa.sharedSizeBytes - 1024;
    const int cand = (int)(std::min<long long>(share, 1 << 20) / 1024 * 1024);
    // a refused attribute leaves the default size (not an engine failure)
    bool allowed = cand <= E->lds_max[w];
    if (cand > bytes && !allowed &&
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, cand) == hipSuccess) {
      E->lds_max[w] = cand;
      allowed = true;
    }
    if (cand > bytes && allowed) {
      int per_cu2 = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, fn, T, cand) == hipSuccess && per_cu2 >= per_cu)
        bytes = cand;
    }
  }
  E->lds_T[w] = T;
  E->lds_B[w] = bytes;
  return bytes;
}

bool ensure_graph(Engine* E, Bucket& B, int T, int tpc) {
  if (B.exec && B.g_T == T && B.g_tpc == tpc) return true;
  if (B.exec) hipGraphExecDestroy(B.exec);
  if (B.graph) hipGraphDestroy(B.graph);
  B.exec = nullptr;
  B.graph = nullptr;
  const bool sliced = E->bits == 32 && B.sliced;
  const int lds = team_lds_bytes(E, T, sliced);
  if (lds < 0) return false;
  if (!ok(E, hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal), "begin capture")) return false;
  const dim3 grid((unsigned)(tpc * E->cus)), block((unsigned)T);
  if (E->bits != 32)
    hipLaunchKernelGGL((k_epoch<double, false>), grid, block, (unsigned)lds, B.stream, B.tab, B.counters, lds);
  else if (sliced)
    hipLaunchKernelGGL((k_epoch<float, true>), grid, block, (unsigned)lds, B.stream, B.tab, B.counters, lds);
  else
    hipLaunchKernelGGL((k_epoch<float, false>), grid, block, (unsigned)lds, B.stream, B.tab, B.counters, lds);
  if (!ok(E, hipStreamEndCapture(B.stream, &B.graph), "end capture")) return false;
  if (!ok(E, hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0), "graph instantiate")) return false;
  B.g_T = T;
  B.g_tpc = tpc;
  return true;
}

constexpr int kAddThreads = 8;  // worker threads of osqp_batch_add_many

// Validates a problem (src/auxil.c validate_data) and builds its host image.
// Reads the engine's fixed configuration only, so problems may be prepared
// on several threads at once.  Returns 0 with *out set, or an error code.
// dev: any_device(...) of the problem's arrays
int prepare(Engine* E, const OSQPCscMatrix* P, const OSQPFloat* q, const OSQPCscMatrix* A, const OSQPFloat* l,
            const OSQPFloat* u, const OSQPFloat* x0, const OSQPFloat* y0, OSQPInt resident, bool dev,
            HProb** out) {
  *out = nullptr;
  if (!P || !A || !q) return OSQP_DATA_VALIDATION_ERROR;
  HostInputs H;
  if (!host_inputs(P, q, A, l, u, x0, y0, dev, &H)) return OSQP_DATA_VALIDATION_ERROR;
  P = &H.P;
  A = &H.A;
  q = H.q;
  l = H.l;
  u = H.u;
  x0 = H.x0;
  y0 = H.y0;
  const int n = P->n, m = A->m;
  if (n <= 0 || m < 0 || P->m != n || A->n != n) return OSQP_DATA_VALIDATION_ERROR;
  if (m > 0 && (!l || !u)) return OSQP_DATA_VALIDATION_ERROR;
  if (n > kMaxRedBlocks * kRedBlock || m > kMaxRedBlocks * kRedBlock) {
    fprintf(stderr, "osqp_batch: problem too large for the batched engine (n, m <= %d)\n", kMaxRedBlocks * kRedBlock);
    return OSQP_DATA_VALIDATION_ERROR;
  }
  for (int j = 0; j < n; ++j)
    for (int k = P->p[j]; k < P->p[j + 1]; ++k)
      if (P->i[k] > j || P->i[k] < 0) return OSQP_DATA_VALIDATION_ERROR;
  for (int j = 0; j < n; ++j)
    for (int k = A->p[j]; k < A->p[j + 1]; ++k)
      if (A->i[k] < 0 || A->i[k] >= m) return OSQP_DATA_VALIDATION_ERROR;
  for (int i = 0; i < m; ++i)
    if (l[i] > u[i]) return OSQP_DATA_VALIDATION_ERROR;
  HProb* p = new HProb();
  p->resident = resident ? 1 : 0;
  p->L = build_host(&p->h, E->bits / 8, P, q, A, l, u, x0, y0, resident != 0, &E->images);
  if (p->L.total > E->arena_size || p->L.st_total + p->L.rs_total > E->pool_size) {
    // could never be admitted: larger than the whole arena or staging pool
    fprintf(stderr, "osqp_batch: problem needs %zu bytes of device arena (of %zu) and %zu of staging (of %zu)\n",
            p->L.total, E->arena_size, p->L.st_total + p->L.rs_total, E->pool_size);
    if (p->h.img_pinned) E->images.give(p->h.img_off);
    delete p;
    return OSQP_DATA_VALIDATION_ERROR;
  }
  p->bucket = bucket_of(p->h);
  *out = p;
  return 0;
}

// Queues a prepared problem (in arrival order); returns its id.
int enqueue(Engine* E, HProb* p) {
  p->id = E->next_id++;
  p->state = kQueued;
  p->t_req = now_s();
  E->probs[p->id] = p;
  E->queue.push_back(p->id);
  return p->id;
}

}  // namespace
}  // namespace osqp_serve

using namespace osqp_serve;

struct OSQPBatch {
  Engine e;
};

extern "C" {

OSQPInt osqp_batch_create(OSQPBatch** batch, const OSQPSettings* s, OSQPInt precision, OSQPInt max_active,
                          long long arena_bytes) {
  if (!batch || !s) return OSQP_SETTINGS_VALIDATION_ERROR;
  *batch = nullptr;
  if (precision != 32 && precision != 64) return OSQP_SETTINGS_VALIDATION_ERROR;
  if (s->rho_is_vec != 1 || s->polishing != 0) {
    fprintf(stderr, "osqp_batch: rho_is_vec = 1 and polishing = 0 are required\n");
    return OSQP_SETTINGS_VALIDATION_ERROR;
  }
  if (s->adaptive_rho < OSQP_ADAPTIVE_RHO_UPDATE_DISABLED || s->adaptive_rho > OSQP_ADAPTIVE_RHO_UPDATE_KKT_ERROR) {
    fprintf(stderr, "osqp_batch: adaptive_rho must be 0, 1, 2 or 3\n");
    return OSQP_SETTINGS_VALIDATION_ERROR;
  }
  if (s->time_limit <= 0 || s->adaptive_rho_fraction <= 0) {
    fprintf(stderr, "osqp_batch: time_limit and adaptive_rho_fraction must be positive\n");
    return OSQP_SETTINGS_VALIDATION_ERROR;
  }
  if (s->rho <= 0 || s->sigma <= 0 || s->alpha <= 0 || s->alpha >= 2 || s->max_iter <= 0 || s->eps_abs < 0 ||
      s->eps_rel < 0 || (s->eps_abs == 0 && s->eps_rel == 0) || s->eps_prim_inf <= 0 || s->eps_dual_inf <= 0 ||
      s->scaling < 0 || s->check_termination < 0 || s->cg_max_iter <= 0 || s->cg_tol_reduction <= 0 ||
      s->cg_tol_fraction <= 0 || s->cg_tol_fraction >= 1 || s->adaptive_rho_tolerance < 1.0 ||
      s->adaptive_rho_interval < 0) {
    fprintf(stderr, "osqp_batch: invalid settings\n");
    return OSQP_SETTINGS_VALIDATION_ERROR;
  }
  if (osqp_algebra_init_libs(s->device)) return OSQP_MEM_ALLOC_ERROR;
  // time_limit in ticks of the device wall clock (wall_clock64); limits of a
  // million seconds or more (the default is 1e10 s) are treated as none, so
  // the default loop reads no clock
  long long tl_ticks = 0;
  if (s->time_limit < 1e6) {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device) != hipSuccess || khz <= 0) {
      fprintf(stderr, "osqp_batch: no device wall clock for time_limit\n");
      osqp_algebra_free_libs();
      return OSQP_SETTINGS_VALIDATION_ERROR;
    }
    tl_ticks = (long long)(s->time_limit * 1e3 * (double)khz);
    if (tl_ticks < 1) tl_ticks = 1;
  }
  OSQPBatch* b = new OSQPBatch();
  Engine* E = &b->e;
  BSettings& B = E->set;
  B.rho0 = s->rho;
  B.sigma = s->sigma;
  B.alpha = s->alpha;
  B.eps_abs = s->eps_abs;
  B.eps_rel = s->eps_rel;
  B.eps_pinf = s->eps_prim_inf;
  B.eps_dinf = s->eps_dual_inf;
  B.ad_tol = s->adaptive_rho_tolerance;
  B.cg_tol_fraction = s->cg_tol_fraction;
  B.max_iter = s->max_iter;
  B.scaling = s->scaling;
  B.check_term = s->check_termination;
  B.adaptive = (int)s->adaptive_rho;
  B.ad_interval = s->adaptive_rho_interval;
  B.ad_fraction = s->adaptive_rho_fraction;
  // osqp_setup's default intervals (src/osqp_api.c:607-620); the time mode
  // fixes its interval during the first solve
  if (B.adaptive == OSQP_ADAPTIVE_RHO_UPDATE_ITERATIONS && B.ad_interval == 0)
    B.ad_interval = s->check_termination ? OSQP_ADAPTIVE_RHO_MULTIPLE_TERMINATION * s->check_termination
                                         : OSQP_ADAPTIVE_RHO_FIXED;
  if (B.adaptive == OSQP_ADAPTIVE_RHO_UPDATE_KKT_ERROR && B.ad_interval == 0) B.ad_interval = 1;
  B.scaled_term = s->scaled_termination;
  B.check_gap = s->check_dualgap;
  B.warm = s->warm_starting;
  B.cg_max_iter = s->cg_max_iter;
  B.cg_red_thresh = s->cg_tol_reduction;
  E->bits = (int)precision;
  E->max_active = max_active > 0 ? (int)max_active : 4096;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, s->device) == hipSuccess && prop.multiProcessorCount > 0)
    E->cus = prop.multiProcessorCount;
  if (hipGetDeviceProperties(&prop, s->device) == hipSuccess && prop.maxSharedMemoryPerMultiProcessor > 0)
    E->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
  B.tl_ticks = tl_ticks;
  bool good = true;
  E->arena_size = arena_bytes > 0 ? (size_t)arena_bytes : ((size_t)4 << 30);
  good = good && ok(E, hipMalloc((void**)&E->arena, E->arena_size), "hipMalloc arena");
  E->pool_size = std::max<size_t>(E->arena_size / 4, (size_t)64 << 20);
  E->pool_size = std::min<size_t>(E->pool_size, (size_t)1 << 30);
  good = good && ok(E, hipHostMalloc((void**)&E->pool, E->pool_size, hipHostMallocDefault), "hipHostMalloc pool");
  // the image arena: half the pool's size; without it images go through the heap
  if (good) {
    E->images.size = E->pool_size / 2;
    if (hipHostMalloc((void**)&E->images.base, E->images.size, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      E->images.base = nullptr;
      E->images.size = 0;
    }
    E->images.alloc.reset(E->images.size);
  }
  for (int k = 0; k < kBuckets && good; ++k) {
    Bucket& Bk = E->buckets[k];
    Bk.sliced = k >= kSlicedBucket;
    good = ok(E, hipStreamCreateWithFlags(&Bk.stream, hipStreamNonBlocking), "stream") &&
           ok(E, hipMalloc((void**)&Bk.counters, sizeof(int) * kCounterRing), "hipMalloc counters") &&
           ok(E, hipMemset(Bk.counters, 0, sizeof(int) * kCounterRing), "hipMemset counters") &&
           ok(E, hipEventCreate(&Bk.e0), "event") && ok(E, hipEventCreate(&Bk.e1), "event");
  }
  if (!good) {
    osqp_batch_destroy(b);
    return OSQP_MEM_ALLOC_ERROR;
  }
  good = good && ok(E, hipEventCreate(&E->ref), "event") && ok(E, hipEventRecord(E->ref, E->buckets[0].stream), "event") &&
         ok(E, hipEventSynchronize(E->ref), "event");
  if (!good) {
    osqp_batch_destroy(b);
    return OSQP_MEM_ALLOC_ERROR;
  }
  E->arena_alloc.reset(E->arena_size);
  E->pool_alloc.reset(E->pool_size);
  *batch = b;
  return 0;
}

OSQPInt osqp_batch_add(OSQPBatch* b, const OSQPCscMatrix* P, const OSQPFloat* q, const OSQPCscMatrix* A,
                       const OSQPFloat* l, const OSQPFloat* u, const OSQPFloat* x0, const OSQPFloat* y0,
                       OSQPInt resident, OSQPInt* id) {
  if (!b || !id) return OSQP_DATA_VALIDATION_ERROR;
  HProb* p = nullptr;
  const bool dev = P && A && osqp_serve::any_device(P, q, A, l, u, x0, y0);
  const int rc = osqp_serve::prepare(&b->e, P, q, A, l, u, x0, y0, resident, dev, &p);
  if (rc) return rc;
  *id = osqp_serve::enqueue(&b->e, p);
  return 0;
}

OSQPInt osqp_batch_add_many(OSQPBatch* b, OSQPInt count, const OSQPCscMatrix* const* P, const OSQPFloat* const* q,
                            const OSQPCscMatrix* const* A, const OSQPFloat* const* l, const OSQPFloat* const* u,
                            const OSQPFloat* const* x0, const OSQPFloat* const* y0, OSQPInt resident,
                            OSQPInt* ids) {
  if (!b || count < 0 || (count > 0 && (!P || !q || !A || !l || !u || !ids))) return OSQP_DATA_VALIDATION_ERROR;
  std::vector<HProb*> ready((size_t)count, nullptr);
  std::vector<int> codes((size_t)count, 0);
  // which problems bring device arrays: on this thread (the runtime's pointer
  // queries serialise concurrent callers)
  std::vector<char> dev((size_t)count, 0);
  for (int k = 0; k < count; ++k)
    if (P[k] && A[k])
      dev[(size_t)k] = osqp_serve::any_device(P[k], q[k], A[k], l[k], u[k], x0 ? x0[k] : nullptr,
                                              y0 ? y0[k] : nullptr);
  // the host-side preparation (validation, CSR / row-group image) of each
  // problem touches only its own data: spread it over worker threads, then
  // queue in order on this thread
  std::atomic<int> next{0};
  auto work = [&]() {
    for (int k = next.fetch_add(1); k < count; k = next.fetch_add(1))
      codes[(size_t)k] = osqp_serve::prepare(&b->e, P[k], q[k], A[k], l[k], u[k], x0 ? x0[k] : nullptr,
                                             y0 ? y0[k] : nullptr, resident, dev[(size_t)k] != 0,
                                             &ready[(size_t)k]);
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const int nthreads = (int)std::min<long>(std::min<long>(kAddThreads, hw ? (long)hw : 1L), count / 8 + 1);
  std::vector<std::thread> pool;
  try {
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(work);
  } catch (...) {
    // no more threads: the ones started and this one take all the problems
  }
  work();
  for (auto& th : pool) th.join();
  OSQPInt first_error = 0;
  for (int k = 0; k < count; ++k) {
    if (codes[(size_t)k]) {
      ids[k] = -1;
      if (!first_error) first_error = codes[(size_t)k];
    } else {
      ids[k] = osqp_serve::enqueue(&b->e, ready[(size_t)k]);
    }
  }
  return first_error;
}

OSQPInt osqp_batch_poll_many(OSQPBatch* b, OSQPInt count, const OSQPInt* ids, OSQPFloat* x, OSQPFloat* y,
                             OSQPBatchInfo* infos, OSQPInt* states) {
  if (!b || count < 0 || (count > 0 && (!ids || !infos || !states))) return -1;
  Engine* E = &b->e;
  long long ox = 0, oy = 0;
  OSQPInt polled = 0;
  for (OSQPInt k = 0; k < count; ++k) {
    auto it = E->probs.find(ids[k]);
    memset(&infos[k], 0, sizeof(OSQPBatchInfo));
    if (it == E->probs.end()) {
      states[k] = -1;
      continue;
    }
    const HProb* p = it->second;
    if (p->state != kEnded) {
      states[k] = p->state == kIdle ? 2 : 0;
      continue;
    }
    const int n = p->h.n, m = p->h.m;
    states[k] = osqp_batch_poll(b, ids[k], x ? x + ox : nullptr, y ? y + oy : nullptr, nullptr, nullptr, &infos[k]);
    ox += n;
    oy += m;
    polled++;
  }
  return polled;
}

OSQPInt osqp_batch_update_data_mat(OSQPBatch* b, OSQPInt id, const OSQPFloat* Px_new, const OSQPInt* Px_new_idx,
                                   OSQPInt P_new_n, const OSQPFloat* Ax_new, const OSQPInt* Ax_new_idx,
                                   OSQPInt A_new_n) {
  if (!b) return -1;
  Engine* E = &b->e;
  auto it = E->probs.find(id);
  if (it == E->probs.end()) return -1;
  HProb* p = it->second;
  if (!p->resident || p->state != kIdle || !p->has_dev || !p->has_res) return -1;
  const HostData& h = p->h;
  // the argument checks of src/osqp_api.c:1335-1366 (1: P, 2: A), plus
  // index ranges, since the engine writes at the indexed slots
  const int nnzP = h.nnzP_in, nnzA = h.nnzA;
  if (P_new_n > nnzP || P_new_n < 0) return 1;
  if (!Px_new_idx && P_new_n != 0 && P_new_n != nnzP) return 1;
  if (P_new_n == 0) P_new_n = nnzP;
  if (A_new_n > nnzA || A_new_n < 0) return 2;
  if (!Ax_new_idx && A_new_n != 0 && A_new_n != nnzA) return 2;
  if (A_new_n == 0) A_new_n = nnzA;
  std::vector<double> pv, av;
  std::vector<int> pi, ai;
  if (Px_new) {
    pv.resize((size_t)P_new_n);
    copy_in(pv.data(), Px_new, sizeof(double) * pv.size());
    if (Px_new_idx) {
      pi.resize((size_t)P_new_n);
      copy_in(pi.data(), Px_new_idx, sizeof(int) * pi.size());
      for (int k : pi)
        if (k < 0 || k >= nnzP) return 1;
    }
  }
  if (Ax_new) {
    av.resize((size_t)A_new_n);
    copy_in(av.data(), Ax_new, sizeof(double) * av.size());
    if (Ax_new_idx) {
      ai.resize((size_t)A_new_n);
      copy_in(ai.data(), Ax_new_idx, sizeof(int) * ai.size());
      for (int k : ai)
        if (k < 0 || k >= nnzA) return 2;
    }
  }
  // each P triangle entry goes to its slot and, off the diagonal, to the
  // mirrored one; each A entry to its CSR slot and to its A' (CSC) slot
  for (size_t j = 0; j < pv.size(); ++j) {
    const int k = pi.empty() ? (int)j : pi[j];
    p->up_pslot.push_back(h.pslot_lo[(size_t)k]);
    p->up_pval.push_back(pv[j]);
    if (h.pslot_up[(size_t)k] >= 0) {
      p->up_pslot.push_back(h.pslot_up[(size_t)k]);
      p->up_pval.push_back(pv[j]);
    }
  }
  for (size_t j = 0; j < av.size(); ++j) {
    const int k = ai.empty() ? (int)j : ai[j];
    p->up_aslot.push_back(h.aslot[(size_t)k]);
    p->up_tslot.push_back(k);
    p->up_aval.push_back(av[j]);
  }
  return 0;
}

OSQPInt osqp_batch_resolve(OSQPBatch* b, OSQPInt id, const OSQPFloat* q_new) {
  if (!b) return -1;
  Engine* E = &b->e;
  auto it = E->probs.find(id);
  if (it == E->probs.end()) return -1;
  HProb* p = it->second;
  if (!p->resident || p->state != kIdle || !p->has_dev || !p->has_res) return -1;
  p->cmd = 2;
  if (!p->up_pslot.empty() || !p->up_aslot.empty()) {
    // the staged matrix values, as one host-mapped record read by the GPU
    const size_t np = p->up_pslot.size(), na = p->up_aslot.size();
    const size_t o_ps = align_up(sizeof(BMatUpd), 64), o_pv = align_up(o_ps + 4 * np, 64);
    const size_t o_as = align_up(o_pv + 8 * np, 64), o_ts = align_up(o_as + 4 * na, 64);
    const size_t o_av = align_up(o_ts + 4 * na, 64), bytes = o_av + 8 * na;
    size_t off = 0;
    if (!E->pool_alloc.alloc(bytes, &off)) return -2;
    char* rec = E->pool + off;
    memcpy(rec + o_ps, p->up_pslot.data(), 4 * np);
    memcpy(rec + o_pv, p->up_pval.data(), 8 * np);
    memcpy(rec + o_as, p->up_aslot.data(), 4 * na);
    memcpy(rec + o_ts, p->up_tslot.data(), 4 * na);
    memcpy(rec + o_av, p->up_aval.data(), 8 * na);
    BMatUpd* u = (BMatUpd*)rec;
    u->np = (int)np;
    u->na = (int)na;
    u->pslot = (const int*)(rec + o_ps);
    u->pval = (const double*)(rec + o_pv);
    u->aslot = (const int*)(rec + o_as);
    u->tslot = (const int*)(rec + o_ts);
    u->aval = (const double*)(rec + o_av);
    p->up_pslot.clear();
    p->up_pval.clear();
    p->up_aslot.clear();
    p->up_tslot.clear();
    p->up_aval.clear();
    p->upd_off = off;
    p->has_upd = true;
    p->cmd = 3;
  }
  char* rs = E->pool + p->res_off;
  BResult* out = (BResult*)(rs + p->L.rs_res);
  memset(out, 0, sizeof(BResult));
  double* qb = (double*)(rs + p->L.rs_q);
  if (q_new) copy_in(qb, q_new, sizeof(double) * (size_t)p->h.n);
  p->qnew = q_new ? 1 : 0;
  p->state = kActive;
  p->t_req = now_s();
  // joins its bucket at the next launch (an epoch may be in flight)
  E->resumed.push_back(p->id);
  E->n_active++;
  return 0;
}

}  // extern "C"

namespace osqp_serve {
namespace {

// Ends bucket k's epoch (finished on the GPU): its results, the tuner, the
// counters.  Returns the number of solves that ended in it.
int finish_bucket(Engine* E, int k) {
  Bucket& B = E->buckets[k];
  B.inflight = false;
  float ms = 0.f, t0 = 0.f, t1 = 0.f;
  hipEventElapsedTime(&ms, B.e0, B.e1);
  // GPU busy time: the union of the buckets' epoch intervals (approximate
  // where intervals are finished out of order)
  if (hipEventElapsedTime(&t0, E->ref, B.e0) == hipSuccess && hipEventElapsedTime(&t1, E->ref, B.e1) == hipSuccess) {
    const double a = std::max((double)t0, E->busy_until_ms);
    if (t1 > a) E->gpu_s += (t1 - a) * 1e-3;
    E->busy_until_ms = std::max(E->busy_until_ms, (double)t1);
  }
  const double t_now = now_s();
  int done = 0;
  long long its = 0;
  std::vector<int> still;
  still.reserve(B.active.size());
  for (size_t i = 0; i < B.active.size(); ++i) {
    HProb* p = E->probs[B.active[i]];
    its += B.tab->e[i].iters_done;
    p->iters_since_add += B.tab->e[i].iters_done;
    // setup done: the stage is no longer read
    if (p->cmd == 1) free_problem_memory(E, p, false, true, false);
    if (p->cmd == 3 && p->has_upd) {
      E->pool_alloc.release(p->upd_off);
      p->has_upd = false;
    }
    p->cmd = 0;
    const BResult* out = (const BResult*)(E->pool + p->res_off + p->L.rs_res);
    if (out->done) {
      p->state = kEnded;
      p->latency = t_now - p->t_req;
      E->n_active--;
      E->last_done.push_back(p->id);
      done++;
    } else {
      still.push_back(p->id);
    }
  }
  B.active.swap(still);
  E->admm_iters += its;
  E->epochs++;
  // online tuner: iterations per second of each candidate on a full epoch
  if (B.trial < trials_of(k)) {
    B.rate[cand_of_trial(k, B.trial)] += ms > 0 ? its / (double)ms : 0.0;
    B.trial++;
    if (B.trial == trials_of(k)) {
      int best = 0;
      for (int c = 1; c < cands_of(k); ++c)
        if (B.rate[c] > B.rate[best]) best = c;
      B.best = best;
    }
  }
  return done;
}

// Waits until at least one bucket in flight has ended (none in flight: no
// wait) and finishes every bucket that has.  Returns the solves that ended,
// or -1 on a HIP error.
int finish_ready(Engine* E) {
  for (;;) {
    bool any = false, ready[kBuckets] = {};
    int n_ready = 0;
    for (int k = 0; k < kBuckets; ++k) {
      Bucket& B = E->buckets[k];
      if (!B.inflight) continue;
      any = true;
      const hipError_t q = hipEventQuery(B.e1);
      if (q == hipSuccess) {
        ready[k] = true;
        n_ready++;
      } else if (q != hipErrorNotReady) {
        ok(E, q, "epoch");
        return -1;
      }
    }
    if (!any) return 0;
    if (n_ready) {
      int done = 0;
      for (int k = 0; k < kBuckets; ++k)
        if (ready[k]) done += finish_bucket(E, k);
      return done;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

}  // namespace
}  // namespace osqp_serve

extern "C" {

// One step of the pipelined engine.  Each size bucket runs its epochs on its
// own stream, independently of the others.  A step admits queued problems
// (their memory reserved and images written) while epochs are in flight,
// waits until at least one bucket's epoch has ended (none in flight: no
// wait), finishes every bucket whose epoch has ended, launches the next epoch
// of every bucket that is idle and has work, and returns without waiting for
// them.  So the host work between two calls (adds, polls, resolves, the
// caller's own) overlaps the GPU, and a bucket of small problems does not
// wait for the epochs of a bucket of large ones.
OSQPInt osqp_batch_step(OSQPBatch* b, OSQPInt* completed) {
  if (completed) *completed = 0;
  if (!b) return OSQP_MEM_ALLOC_ERROR;
  Engine* E = &b->e;
  if (E->failed) return OSQP_MEM_ALLOC_ERROR;
  E->last_done.clear();
  // admission in arrival order while slots and memory allow (the tables of
  // the epoch in flight are not touched)
  std::vector<int> fresh;
  size_t kept = 0;
  for (size_t k = 0; k < E->queue.size(); ++k) {
    HProb* p = E->probs[E->queue[k]];
    if (E->n_active + (int)fresh.size() < E->max_active && admit(E, p)) {
      fresh.push_back(p->id);
    } else {
      E->queue[kept++] = E->queue[k];
    }
  }
  E->queue.resize(kept);
  // the admitted problems' images into their stages (while the epochs in
  // flight run), before any of them is launched
  flush_stage_copies(E);
  const int done = finish_ready(E);
  if (done < 0) return OSQP_MEM_ALLOC_ERROR;
  for (int id : fresh) {
    HProb* p = E->probs[id];
    p->state = kActive;
    E->buckets[p->bucket].incoming.push_back(id);
    E->n_active++;
  }
  for (int id : E->resumed) E->buckets[E->probs[id]->bucket].incoming.push_back(id);
  E->resumed.clear();
  // launch every idle bucket with work (a bucket's table and active list are
  // left alone while its epoch is in flight)
  for (int k = 0; k < kBuckets; ++k) {
    Bucket& B = E->buckets[k];
    if (B.inflight) continue;
    B.active.insert(B.active.end(), B.incoming.begin(), B.incoming.end());
    B.incoming.clear();
    if (B.active.empty()) continue;
    if (!ensure_table(E, B, B.active.size())) return OSQP_MEM_ALLOC_ERROR;
    B.tab->count = (int)B.active.size();
    B.tab->epoch = ++B.epoch;
    B.tab->iters = (long long)B.active.size() > 2LL * E->cus ? kEpochIters : kEpochItersLight;
    for (size_t i = 0; i < B.active.size(); ++i) {
      HProb* p = E->probs[B.active[i]];
      BEntry& en = B.tab->e[i];
      en.dev = (BProb*)(E->arena + p->dev_off);
      en.init = p->cmd == 1 ? (const BProb*)(E->pool + p->stage_off) : nullptr;
      en.cmd = p->cmd;
      en.iters_done = 0;
      en.upd_q = nullptr;
      if ((p->cmd == 2 || p->cmd == 3) && p->qnew) en.upd_q = (const double*)(E->pool + p->res_off + p->L.rs_q);
      en.upd_mat = p->cmd == 3 ? (const BMatUpd*)(E->pool + p->upd_off) : nullptr;
    }
    const int c = B.trial < trials_of(k) ? cand_of_trial(k, B.trial) : B.best;
    const int T = B.fixT ? B.fixT : kCands[c].T, tpc = B.fixT ? B.fixTpc : kCands[c].tpc;
    if (!ensure_graph(E, B, T, tpc)) return OSQP_MEM_ALLOC_ERROR;
    if (!ok(E, hipEventRecord(B.e0, B.stream), "event record")) return OSQP_MEM_ALLOC_ERROR;
    if (!ok(E, hipGraphLaunch(B.exec, B.stream), "graph launch")) return OSQP_MEM_ALLOC_ERROR;
    if (!ok(E, hipEventRecord(B.e1, B.stream), "event record")) return OSQP_MEM_ALLOC_ERROR;
    B.inflight = true;
  }
  if (completed) *completed = done;
  return 0;
}

OSQPInt osqp_batch_last_completed(const OSQPBatch* b, OSQPInt* ids, OSQPInt cap) {
  if (!b) return 0;
  const std::vector<int>& d = b->e.last_done;
  if (ids)
    for (size_t k = 0; k < d.size() && (OSQPInt)k < cap; ++k) ids[k] = d[k];
  return (OSQPInt)d.size();
}

OSQPInt osqp_batch_poll(OSQPBatch* b, OSQPInt id, OSQPFloat* x, OSQPFloat* y, OSQPFloat* prim_inf_cert,
                        OSQPFloat* dual_inf_cert, OSQPBatchInfo* info) {
  if (!b) return -1;
  Engine* E = &b->e;
  auto it = E->probs.find(id);
  if (it == E->probs.end()) return -1;
  HProb* p = it->second;
  if (p->state != kEnded) return p->state == kIdle ? 1 : 0;
  const char* rs = E->pool + p->res_off;
  const BResult* out = (const BResult*)(rs + p->L.rs_res);
  if (x) memcpy(x, rs + p->L.rs_x, sizeof(double) * (size_t)p->h.n);
  if (y && p->h.m) memcpy(y, rs + p->L.rs_y, sizeof(double) * (size_t)p->h.m);
  if (prim_inf_cert && p->h.m) memcpy(prim_inf_cert, rs + p->L.rs_pc, sizeof(double) * (size_t)p->h.m);
  if (dual_inf_cert) memcpy(dual_inf_cert, rs + p->L.rs_dc, sizeof(double) * (size_t)p->h.n);
  if (info) {
    info->status_val = out->status;
    info->iter = out->iter;
    info->rho_updates = out->rho_updates;
    info->obj_val = out->obj_val;
    info->dual_obj_val = out->dual_obj_val;
    info->prim_res = out->prim_res;
    info->dual_res = out->dual_res;
    info->duality_gap = out->gap;
    info->rho_estimate = out->rho_estimate;
    info->primdual_int = out->primdual_int;
    info->rel_kkt_error = out->rel_kkt;
    info->latency = p->latency;
    info->cg_iters = out->cg_iters;
  }
  if (p->resident) {
    p->state = kIdle;
  } else {
    free_problem_memory(E, p, true, true, true);
    E->probs.erase(it);
    delete p;
  }
  return 1;
}

OSQPInt osqp_batch_release(OSQPBatch* b, OSQPInt id) {
  if (!b) return -1;
  Engine* E = &b->e;
  auto it = E->probs.find(id);
  if (it == E->probs.end()) return -1;
  HProb* p = it->second;
  if (p->state == kActive) return -1;
  if (p->state == kQueued) E->queue.erase(std::remove(E->queue.begin(), E->queue.end(), p->id), E->queue.end());
  free_problem_memory(E, p, true, true, true);
  E->probs.erase(it);
  delete p;
  return 0;
}

OSQPInt osqp_batch_pending(const OSQPBatch* b) {
  if (!b) return 0;
  return (OSQPInt)(b->e.queue.size() + (size_t)b->e.n_active);
}

OSQPInt osqp_batch_tuning(const OSQPBatch* b, OSQPInt bucket, OSQPInt* team, OSQPInt* tpc) {
  if (!b || bucket < 0 || bucket >= kBuckets) return -1;
  const Bucket& B = b->e.buckets[bucket];
  if (B.trial < trials_of(bucket)) return -1;
  if (team) *team = B.fixT ? B.fixT : kCands[B.best].T;
  if (tpc) *tpc = B.fixT ? B.fixTpc : kCands[B.best].tpc;
  return 0;
}

OSQPInt osqp_batch_set_team(OSQPBatch* b, OSQPInt bucket, OSQPInt team, OSQPInt tpc) {
  if (!b || bucket < 0 || bucket >= kBuckets || tpc < 1 || tpc > 64) return -1;
  if (team != 64 && team != 128 && team != 256 && team != 512) return -1;
  Bucket& B = b->e.buckets[bucket];
  B.fixT = (int)team;
  B.fixTpc = (int)tpc;
  B.trial = trials_of(bucket);
  return 0;
}

OSQPInt osqp_batch_counters(const OSQPBatch* b, long long* epochs, long long* iters, OSQPFloat* gpu_s) {
  if (!b) return -1;
  if (epochs) *epochs = b->e.epochs;
  if (iters) *iters = b->e.admm_iters;
  if (gpu_s) *gpu_s = b->e.gpu_s;
  return 0;
}

OSQPInt osqp_batch_destroy(OSQPBatch* b) {
  if (!b) return 0;
  Engine* E = &b->e;
  for (int k = 0; k < kBuckets; ++k)
    if (E->buckets[k].stream) hipStreamSynchronize(E->buckets[k].stream);
  for (auto& kv : E->probs) delete kv.second;
  E->probs.clear();
  for (int k = 0; k < kBuckets; ++k) {
    Bucket& B = E->buckets[k];
    if (B.exec) hipGraphExecDestroy(B.exec);
    if (B.graph) hipGraphDestroy(B.graph);
    if (B.tab) hipHostFree(B.tab);
    if (B.counters) hipFree(B.counters);
    if (B.e0) hipEventDestroy(B.e0);
    if (B.e1) hipEventDestroy(B.e1);
    if (B.stream) hipStreamDestroy(B.stream);
  }
  if (E->ref) hipEventDestroy(E->ref);
  if (E->arena) hipFree(E->arena);
  if (E->pool) hipHostFree(E->pool);
  if (E->images.base) hipHostFree(E->images.base);
  delete b;
  osqp_algebra_free_libs();
  return 0;
}

}  // extern "C"
