// CPU check of the batched engine's matrix layout (tests/test_layout_cpu.py
// builds and runs it; no GPU is used).  For every problem of a file of CSC
// problems it builds the stage image with the library's own host code
// (build_host in osqp_amd/csrc/serve/batch.hip), replays the setup kernel's
// derivation of the row groups' columns and values (Solver::fill_group_values)
// and the product loops of team_rows lane by lane for every team width, with a
// bounds check on every index, and checks that
//   * each row is reduced exactly once by the FP32 CG's walk (short-row groups,
//     slices, long and empty rows) and by the FP64 walk (all groups);
//   * every slice entry is the right CSR entry, padding is zero, every slice
//     slot is written once, the slice sums equal the CSR row sums;
//   * group slots hold consecutive CSR entries of one row; long-row padding
//     and chunk loads stay inside the compact copy;
//   * only problems of the sliced buckets get slices;
//   * the team-size tuner times every candidate of a bucket twice, forward
//     then backward, and the small buckets have no 512-thread candidates;
//   * images built in the image arena (here over ordinary memory) equal the
//     heap-built ones, a full arena falls back to the heap, freed blocks
//     coalesce, and concurrent take / give never hand out overlapping blocks.
// usage: layout_check <problems file> [plant]   ("plant" corrupts one slice
// length in each sliced image, which must then fail)
#include "serve/batch.hip"
#include <set>
#include <array>
#include <random>
#include <thread>
using namespace osqp_serve;

static int fails = 0;
static long long n_sliced = 0, n_wide = 0, n_long = 0;  // what the problems exercised
#define CHECK(c, ...) do { if (!(c)) { if (fails++ < 20) { fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); fprintf(stderr, __VA_ARGS__); fprintf(stderr, "\n"); } } } while (0)

static void check_matrix(const char* name, const char* blk, size_t total, int rows, int cols, const GroupDims& d,
                         const Layout::Groups& o, size_t Orp, size_t Oci, size_t Oval, bool fp32) {
  const int* rp = (const int*)(blk + Orp);
  const int* ci = (const int*)(blk + Oci);
  const double* val = (const double*)(blk + Oval);
  const int* grow = (const int*)(blk + o.grow);
  const unsigned char* goff = (const unsigned char*)(blk + o.goff);
  // the setup kernel's derivation (fill_group_values), replayed
  std::vector<int> gk((size_t)d.ng * kGroupSlots, -1);
  for (size_t i = 0; i < gk.size(); ++i) {
    const int r = grow[i];
    if (r < 0) continue;
    CHECK(r < rows && goff[i] < rp[r + 1] - rp[r], "%s: goff", name);
    gk[i] = rp[r] + goff[i];
    // entries of a row sit in consecutive slots of one group, in order
    if (i % kGroupSlots && grow[i - 1] == r) CHECK(gk[i] == gk[i - 1] + 1, "%s: slot order", name);
  }
  const int* longs = (const int*)(blk + o.longs);
  const int* empty = (const int*)(blk + o.empty);
  if (!fp32) CHECK(!d.ell, "%s: slices in the FP64 engine", name);
  std::vector<float> x(cols);
  for (int c = 0; c < cols; ++c) x[c] = (float)std::sin(0.37 * c + 1.0);
  for (int walk = 0; walk < 2; ++walk) {
    const bool w32 = walk == 0 && d.ell;
    if (walk == 0 && !d.ell) continue;
    std::vector<int> seen(rows, 0);
    const int ng = w32 ? d.ng_short : d.ng;
    CHECK(ng <= d.ng, "ng");
    for (long long sidx = 0; sidx < (long long)ng * kGroupSlots; ++sidx) {
      const int r = grow[sidx];
      if (r < 0) continue;
      CHECK(r < rows, "%s: group row %d", name, r);
      const int k = gk[sidx];
      CHECK(k >= rp[r] && k < rp[r + 1], "%s: gk", name);
      if (k == rp[r]) seen[r]++;
      if (w32) CHECK(rp[r + 1] - rp[r] < kEllMin, "%s: medium row %d in a short group", name, r);
    }
    for (int j = 0; j < d.nempty; ++j) seen[empty[j]]++;
    for (int j = 0; j < d.nlong; ++j) seen[longs[j]]++;
    if (w32) {
      const int* eoff = (const int*)(blk + o.eoff);
      const int* erl = (const int*)(blk + o.erl);
      CHECK(o.ev32 + 4 * (size_t)d.ell_slots <= total && o.eci16 + 2 * (size_t)d.ell_slots <= total,
            "%s: slice arrays past the block", name);
      CHECK(o.ev32 % 16 == 0 && o.eci16 % 8 == 0 && o.erl % 8 == 0, "%s: alignment", name);
      const int ne = eoff[d.nslice];
      CHECK(ne == d.ell_slots, "%s: eoff end %d vs %lld", name, ne, d.ell_slots);
      // the setup kernel's derivation, replayed: every slot written once
      std::vector<float> ev32(d.ell_slots, -7.f);
      std::vector<unsigned short> eci16(d.ell_slots, 0xBEEF);
      std::vector<int> esrc(d.ell_slots, -2), wrote(d.ell_slots, 0);
      for (int w = 0; w < d.nslice * kWave; ++w) {
        const int sl = w / kWave, lane = w % kWave;
        const int r = erl[2 * w], len = erl[2 * w + 1];
        const int b0 = eoff[sl], width = (eoff[sl + 1] - b0) / kWave;
        CHECK((eoff[sl + 1] - b0) % kWave == 0, "%s: slice size", name);
        for (int q = 0; q < width; ++q) {
          const int at = b0 + (q / kEllVec) * kWave * kEllVec + lane * kEllVec + q % kEllVec;
          CHECK(at >= 0 && at < d.ell_slots, "%s: fill index %d", name, at);
          wrote[at]++;
          eci16[at] = q < len ? (unsigned short)ci[rp[r] + q] : 0;
          ev32[at] = q < len ? (float)val[rp[r] + q] : 0.f;
          esrc[at] = q < len ? rp[r] + q : -1;
        }
      }
      for (int i = 0; i < ne; ++i) CHECK(wrote[i] == 1, "%s: slice slot %d written %d times", name, i, wrote[i]);
      constexpr int kChunk = kWave * kEllVec;
      for (int W : {1, 2, 4, 8}) {
        std::vector<int> sdone(d.nslice, 0);
        for (int wv = 0; wv < W; ++wv)
          for (int base = wv; base < d.nslice; base += kWave * W) {
            const int cnt = std::min(kWave, (d.nslice - base + W - 1) / W);
            int o0[kWave], o1[kWave];
            for (int lane = 0; lane < kWave; ++lane) {
              const int mine = base + lane * W;
              CHECK(mine >= d.nslice || mine + 1 <= d.nslice, "eoff index");
              o0[lane] = mine < d.nslice ? eoff[mine] : 0;
              o1[lane] = mine < d.nslice ? eoff[mine + 1] : 0;
            }
            for (int i = 0; i < cnt; ++i) {
              const int sl = base + i * W;
              CHECK(sl < d.nslice, "%s: slice %d of %d", name, sl, d.nslice);
              sdone[sl]++;
              const int b0 = o0[i], b1 = o1[i], nch = (b1 - b0) / kChunk;
              CHECK(b0 == eoff[sl] && b1 == eoff[sl + 1], "%s: lane-held offsets", name);
              CHECK((b1 - b0) % kChunk == 0 && b0 % kChunk == 0, "%s: slice %d size", name, sl);
              if (W != 8) continue;  // the sums once
              int maxlen = 0;
              for (int lane = 0; lane < kWave; ++lane) {
                const size_t at = (size_t)sl * kWave + lane;
                CHECK(at < (size_t)d.nslice * kWave, "erl index");
                const int r = erl[2 * at], len = erl[2 * at + 1];
                CHECK(r >= -1 && r < rows, "%s: erl row", name);
                CHECK(r < 0 ? len == 0 : len == rp[r + 1] - rp[r], "%s: erl length", name);
                if (r >= 0) {
                  CHECK(len >= kEllMin && len <= kLongRow, "%s: slice row length %d", name, len);
                  seen[r]++;
                }
                maxlen = std::max(maxlen, len);
                float acc = 0.f, ref = 0.f;
                auto add4 = [&](int k, int e0, bool real) {
                  CHECK(k >= 0 && k + kEllVec <= d.ell_slots, "%s: load %d of %lld", name, k, d.ell_slots);
                  for (int q = 0; q < kEllVec; ++q) {
                    const int c = eci16[k + q];
                    CHECK(c < cols, "%s: column", name);
                    const float pr = ev32[k + q] * x[c];
                    if (e0 + q < len) {
                      CHECK(real, "%s: clamped load used", name);
                      CHECK(esrc[k + q] == rp[r] + e0 + q && c == ci[rp[r] + e0 + q], "%s: entry of row %d", name, r);
                      acc += pr;
                    } else if (real) {
                      CHECK(esrc[k + q] == -1 && ev32[k + q] == 0.f, "%s: padding", name);
                    }
                  }
                };
                for (int ii = 0; ii < kEllRounds; ++ii)
                  add4(b0 + (ii < nch ? ii : 0) * kChunk + lane * kEllVec, ii * kEllVec, ii < nch);
                for (int j = kEllRounds; j < nch; ++j) add4(b0 + j * kChunk + lane * kEllVec, j * kEllVec, true);
                if (r >= 0) {
                  for (int kk = rp[r]; kk < rp[r + 1]; ++kk) ref += (float)val[kk] * x[ci[kk]];
                  CHECK(acc == ref, "%s: row %d sum", name, r);
                }
              }
              CHECK(nch * kEllVec == ell_width(maxlen), "%s: slice %d width", name, sl);
              if (nch > kEllRounds) n_wide++;
            }
          }
        for (int sl = 0; sl < d.nslice; ++sl) CHECK(sdone[sl] == 1, "%s: W %d slice %d taken %d times", name, W, sl, sdone[sl]);
      }
    }
    if (w32 && o.compact) {
      // compact long rows: chunks that start past the row's end are not loaded
      const int* loff = (const int*)(blk + o.loff);
      // the setup kernel's derivation, replayed
      std::vector<int> lsrc(d.long_slots, -2);
      for (int j = 0; j < d.nlong; ++j) {
        const int r = longs[j], s0 = loff[j], len = rp[r + 1] - rp[r], pad = loff[j + 1] - s0;
        CHECK(pad >= len && pad < len + kLongAlign && s0 % kLongAlign == 0, "%s: long padding", name);
        for (int q = 0; q < pad; ++q) {
          CHECK(s0 + q < d.long_slots && lsrc[s0 + q] == -2, "%s: long fill", name);
          lsrc[s0 + q] = q < len ? rp[r] + q : -1;
        }
      }
      constexpr int kChunk = kWave * kLongVec;
      for (int j = 0; j < d.nlong; ++j) {
        const int r = longs[j];
        const int s0 = loff[j], e = s0 + (rp[r + 1] - rp[r]);
        std::vector<int> used(e - s0, 0);
        for (int k0 = s0; k0 < e; k0 += kLongVecRounds * kChunk)
          for (int i = 0; i < kLongVecRounds; ++i) {
            if (!(k0 + i * kChunk < e)) continue;
            for (int lane = 0; lane < kWave; ++lane) {
              const int k = k0 + i * kChunk + lane * kLongVec;
              const int kk = k < e ? k : s0;
              CHECK(kk >= 0 && kk + kLongVec <= d.long_slots, "%s: long load %d of %lld", name, kk, d.long_slots);
              CHECK(kk % 4 == 0, "long alignment");
              for (int q = 0; q < kLongVec; ++q)
                if (k + q < e) {
                  CHECK(lsrc[kk + q] == rp[r] + (k + q - s0), "%s: long entry", name);
                  used[k + q - s0]++;
                }
            }
          }
        for (size_t u = 0; u < used.size(); ++u) CHECK(used[u] == 1, "%s: long entry %zu used %d", name, u, used[u]);
      }
    }
    for (int r = 0; r < rows; ++r) CHECK(seen[r] == 1, "%s walk %d: row %d seen %d times (len %d)", name, walk, r, seen[r], rp[r + 1] - rp[r]);
  }
  printf("  %-3s rows %6d ng %5d ng_short %5d long %3d slices %4d ell_slots %8lld%s\n", name, rows, d.ng, d.ng_short,
         d.nlong, d.nslice, d.ell_slots, d.ell ? "" : " (no slices)");
}

// the image arena over ordinary memory: arena-built images equal heap-built
// ones; once full, images go to the heap; giving every block back leaves one
// free range; eight threads taking and giving blocks never overlap
static void check_image_arena(const std::vector<std::array<OSQPCscMatrix, 2>>& probs, const std::vector<int>& ns,
                              const std::vector<int>& ms) {
  size_t total = 0;
  std::vector<double> q, l, u;
  for (size_t t = 0; t < probs.size(); ++t) {
    HostData h;
    q.assign(ns[t], 0.0), l.assign(ms[t], -1.0), u.assign(ms[t], 1.0);
    const Layout L = build_host(&h, 4, &probs[t][0], q.data(), &probs[t][1], l.data(), u.data(), nullptr, nullptr, true);
    total += align_up(L.st_total - L.st_static, 256);
  }
  std::vector<char> mem(total / 2);
  ImageArena ia;
  ia.base = mem.data();
  ia.size = mem.size();
  ia.alloc.reset(ia.size);
  std::vector<std::unique_ptr<HostData>> keep;
  int pinned = 0, heap = 0;
  for (size_t t = 0; t < probs.size(); ++t) {
    q.assign(ns[t], 0.0), l.assign(ms[t], -1.0), u.assign(ms[t], 1.0);
    keep.emplace_back(new HostData());
    HostData& h1 = *keep.back();
    HostData h2;
    const Layout L1 = build_host(&h1, 4, &probs[t][0], q.data(), &probs[t][1], l.data(), u.data(), nullptr, nullptr,
                                 true, &ia);
    build_host(&h2, 4, &probs[t][0], q.data(), &probs[t][1], l.data(), u.data(), nullptr, nullptr, true);
    const size_t bytes = L1.st_total - L1.st_static;
    CHECK(!h2.img_pinned && h2.img_ptr == h2.img.get(), "heap image");
    CHECK(h1.img_pinned == (h1.img_ptr >= mem.data() && h1.img_ptr + bytes <= mem.data() + mem.size()), "arena image");
    // every region the builder writes (the alignment gaps between them are
    // written by no one and read by no one)
    const size_t b0 = L1.s_begin;
    auto same = [&](size_t off, size_t len, const char* what) {
      CHECK(off >= b0 && off - b0 + len <= bytes && memcmp(h1.img_ptr + (off - b0), h2.img_ptr + (off - b0), len) == 0,
            "arena image %zu: %s differs from the heap image", t, what);
    };
    const size_t n = (size_t)ns[t], m = (size_t)ms[t];
    same(L1.Prp, 4 * (n + 1), "Prp");
    same(L1.Pci, 4 * (size_t)h1.nnzP, "Pci");
    same(L1.Pval, 8 * (size_t)h1.nnzP, "Pval");
    same(L1.Pdiag, 4 * n, "Pdiag");
    same(L1.Arp, 4 * (m + 1), "Arp");
    same(L1.Aci, 4 * (size_t)h1.nnzA, "Aci");
    same(L1.Aval, 8 * (size_t)h1.nnzA, "Aval");
    same(L1.Trp, 4 * (n + 1), "Trp");
    same(L1.Tci, 4 * (size_t)h1.nnzA, "Tci");
    same(L1.Tval, 8 * (size_t)h1.nnzA, "Tval");
    same(L1.q, 8 * n, "q");
    same(L1.l, 8 * m, "l");
    same(L1.u, 8 * m, "u");
    const GroupDims* gd[3] = {&h1.gP, &h1.gA, &h1.gT};
    const Layout::Groups* go[3] = {&L1.gP, &L1.gA, &L1.gT};
    for (int k = 0; k < 3; ++k) {
      const GroupDims& g = *gd[k];
      const Layout::Groups& o = *go[k];
      const size_t slots = (size_t)g.ng * kGroupSlots;
      same(o.longs, 4 * (size_t)g.nlong, "longs");
      same(o.empty, 4 * (size_t)g.nempty, "empty");
      same(o.goff, slots, "goff");
      same(o.grow, 4 * slots, "grow");
      if (o.compact) same(o.loff, 4 * ((size_t)g.nlong + 1), "loff");
      if (g.ell) {
        same(o.eoff, 4 * ((size_t)g.nslice + 1), "eoff");
        same(o.erl, 8 * (size_t)g.nslice * kWave, "erl");
      }
    }
    (h1.img_pinned ? pinned : heap)++;
  }
  CHECK(pinned > 0 && heap > 0, "arena of half the images: %d pinned, %d on the heap", pinned, heap);
  for (auto& h : keep)
    if (h->img_pinned) ia.give(h->img_off);
  size_t off = 0;
  CHECK(ia.alloc.alloc(ia.size / 256 * 256, &off) && off == 0, "freed blocks did not coalesce");
  ia.give(off);
  std::atomic<int> overlaps{0};
  auto worker = [&](int id) {
    std::mt19937 rng(id);
    std::vector<std::pair<size_t, size_t>> mine;
    for (int it = 0; it < 2000; ++it) {
      if (mine.size() < 8 && rng() % 2) {
        const size_t b = 256 + rng() % 8192;
        size_t o;
        if (char* pblk = ia.take(b, &o)) {
          memset(pblk, id + 1, b);
          mine.push_back({o, b});
        }
      } else if (!mine.empty()) {
        const auto [o, b] = mine.back();
        mine.pop_back();
        for (size_t k = 0; k < b; ++k)
          if (mem[o + k] != (char)(id + 1)) {
            overlaps++;
            break;
          }
        ia.give(o);
      }
    }
    for (auto [o, b] : mine) ia.give(o);
  };
  std::vector<std::thread> th;
  for (int id = 0; id < 8; ++id) th.emplace_back(worker, id);
  for (auto& x : th) x.join();
  CHECK(overlaps == 0, "concurrent arena blocks overlapped %d times", overlaps.load());
  printf("image arena: %d images pinned, %d on the heap\n", pinned, heap);
}

static void check_tuner_schedule() {
  for (int b = 0; b < kBuckets; ++b) {
    const int nc = cands_of(b);
    CHECK(nc >= 1 && nc <= kNumCands && trials_of(b) == 2 * nc, "bucket %d: %d candidates", b, nc);
    std::vector<int> seen(kNumCands, 0);
    for (int t = 0; t < trials_of(b); ++t) {
      const int c = cand_of_trial(b, t);
      CHECK(c >= 0 && c < nc, "bucket %d trial %d: candidate %d", b, t, c);
      if (c >= 0 && c < kNumCands) seen[c]++;
      CHECK(c == (t < nc ? t : 2 * nc - 1 - t), "bucket %d: trial %d not forward-then-backward", b, t);
    }
    for (int c = 0; c < nc; ++c) CHECK(seen[c] == 2, "bucket %d: candidate %d timed %d times", b, c, seen[c]);
    if (b < kSmallBuckets)
      for (int c = 0; c < nc; ++c) CHECK(kCands[c].T <= 256, "bucket %d tries a %d-thread team", b, kCands[c].T);
  }
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  const bool plant = argc > 2 && strcmp(argv[2], "plant") == 0;
  int np = 0;
  if (fread(&np, 4, 1, f) != 1) return 2;
  std::vector<std::array<OSQPCscMatrix, 2>> all;
  std::vector<int> all_n, all_m;
  std::vector<std::vector<int>> store_i;
  std::vector<std::vector<double>> store_x;
  for (int t = 0; t < np; ++t) {
    int n, m;
    if (fread(&n, 4, 1, f) != 1 || fread(&m, 4, 1, f) != 1) return 2;
    std::vector<int> pp[2], pi[2];
    std::vector<double> px[2];
    OSQPCscMatrix M[2];
    for (int k = 0; k < 2; ++k) {
      int nnz;
      if (fread(&nnz, 4, 1, f) != 1) return 2;
      pp[k].resize(n + 1);
      pi[k].resize(nnz);
      px[k].resize(nnz);
      if (fread(pp[k].data(), 4, n + 1, f) != (size_t)n + 1 || fread(pi[k].data(), 4, nnz, f) != (size_t)nnz ||
          fread(px[k].data(), 8, nnz, f) != (size_t)nnz)
        return 2;
      M[k].m = k == 0 ? n : m;
      M[k].n = n;
      M[k].p = pp[k].data();
      M[k].i = pi[k].data();
      M[k].x = px[k].data();
      M[k].nzmax = M[k].nz = nnz;
      M[k].owned = 0;
    }
    std::vector<double> q(n, 0.0), l(m, -1.0), u(m, 1.0);
    {
      // kept for the arena check
      std::array<OSQPCscMatrix, 2> c = {M[0], M[1]};
      for (int k = 0; k < 2; ++k) {
        store_i.push_back(pp[k]);
        c[k].p = store_i.back().data();
        store_i.push_back(pi[k]);
        c[k].i = store_i.back().data();
        store_x.push_back(px[k]);
        c[k].x = store_x.back().data();
      }
      all.push_back(c);
      all_n.push_back(n);
      all_m.push_back(m);
    }
    for (int vbytes : {4, 8}) {
      HostData h;
      const Layout L = build_host(&h, vbytes, &M[0], q.data(), &M[1], l.data(), u.data(), nullptr, nullptr, true);
      std::vector<char> blk(L.total + 64, 0);
      memcpy(blk.data() + L.s_begin, h.img_ptr, L.s_end - L.s_begin);
      const bool sliced_bucket = size_bucket(n, m, h.nnzA, h.nnzP) >= kSlicedBucket;
      for (const GroupDims* g : {&h.gP, &h.gA, &h.gT}) {
        CHECK(!g->ell || (vbytes == 4 && sliced_bucket), "slices outside the sliced buckets");
        n_sliced += g->ell;
        n_long += g->nlong;
      }
      if (plant && h.gT.ell) ((int*)(blk.data() + L.gT.erl))[1] += 1;
      printf("problem %d n %d m %d vbytes %d total %zu image %zu\n", t, n, m, vbytes, L.total, L.st_total - L.st_static);
      check_matrix("P", blk.data(), L.total, n, n, h.gP, L.gP, L.Prp, L.Pci, L.Pval, vbytes == 4);
      check_matrix("A", blk.data(), L.total, m, n, h.gA, L.gA, L.Arp, L.Aci, L.Aval, vbytes == 4);
      check_matrix("At", blk.data(), L.total, n, m, h.gT, L.gT, L.Trp, L.Tci, L.Tval, vbytes == 4);
    }
  }
  check_image_arena(all, all_n, all_m);
  check_tuner_schedule();
  printf("sliced matrices %lld, slices wider than one load round %lld, long rows %lld\n", n_sliced, n_wide, n_long);
  printf(fails ? "FAILURES: %d\n" : "all checks passed\n", fails);
  return fails ? 1 : 0;
}
