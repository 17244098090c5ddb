// lapgpu_plan.cpp — host preprocessing for the device solve.
//
// The LDLinv sweeps of the reference (forward!/backward!,
// src/approxChol.jl:1369-1408) walk the n-1 columns strictly in elimination
// order.  Written in closed form (own derivation) with, per column c and its
// entries k = 1..K (K the sentinel, fval = 1):
//     w_k = f_k * prod_{l<k} (1 - f_l)   for k < K,     w_K = P_c,
//     P_c = prod_{l<K} (1 - f_l),
// forward! computes y0[v] = b[v] + sum_{columns c that hold row v} w * y0[c]
// and leaves y[c] = P_c y0[c]; the diagonal step divides by d; backward!
// computes yF[c] = P_c * (y[c] + yF[j_K]) + sum_{k<K} w_k yF[j_k].  Hence
//     yF[c] = s_c y0[c] + sum_k w_k yF[j_k],  s_c = P_c^2 / d_c (d_c != 0),
// and the never-eliminated root keeps yF = y0.  Both sweeps become gathers
// ("pull" form): no atomics, a fixed summation order, and each depends only
// on vertices of lower (forward) or higher (backward) DAG level.  Vertices
// are renumbered by forward level so every level is a contiguous range.
#include "lapgpu_plan.hpp"
#include "lapgpu_lock.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>

namespace lapgpu {

namespace {

// Host threads for the planner's parallel loops.  for_each(n, grain, fn)
// calls fn(lo, hi) over [0, n) in chunks of `grain` indices, the calling
// thread taking part; every chunk has finished when it returns.  Each chunk
// writes outputs of its own, so results do not depend on the thread count.
class Team {
 public:
  explicit Team(int threads) {
    for (int t = 1; t < threads; ++t) pool_.emplace_back([this] { work(); });
  }
  ~Team() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& th : pool_) th.join();
  }
  int size() const { return (int)pool_.size() + 1; }
  template <class F>
  void for_each(int64_t n, int64_t grain, F&& fn) {
    if (n <= 0) return;
    if (pool_.empty() || n <= grain) {
      fn(int64_t(0), n);
      return;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      job_ = [&fn](int64_t a, int64_t b) { fn(a, b); };
      n_ = n;
      grain_ = grain;
      next_.store(0);
      busy_ = (int)pool_.size();
      ++gen_;
    }
    cv_.notify_all();
    chunks();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return busy_ == 0; });
    job_ = nullptr;
  }

 private:
  void chunks() {
    for (;;) {
      const int64_t lo = next_.fetch_add(grain_);
      if (lo >= n_) return;
      job_(lo, std::min(n_, lo + grain_));
    }
  }
  void work() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
      }
      chunks();
      std::lock_guard<std::mutex> g(mu_);
      if (--busy_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> pool_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  uint64_t gen_ = 0;
  bool stop_ = false;
  std::function<void(int64_t, int64_t)> job_;
  int64_t n_ = 0, grain_ = 1;
  std::atomic<int64_t> next_{0};
  int busy_ = 0;
};

// Threads for one planning call: LAPGPU_PLAN_THREADS (default: the
// hardware threads, at most 16) shared among the calls running at once.
// A nested call (build_tail inside build_plan) keeps its caller's share.
std::atomic<int> g_active_plans{0};
thread_local int tl_plan_threads = 0;

struct PlanThreads {
  int n;
  bool outer;
  PlanThreads() : outer(tl_plan_threads == 0) {
    if (!outer) {
      n = tl_plan_threads;
      return;
    }
    const int active = ++g_active_plans;
    int base = (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    if (const char* e = getenv("LAPGPU_PLAN_THREADS")) base = std::max(1, std::min(64, atoi(e)));
    n = std::max(1, base / active);
    tl_plan_threads = n;
  }
  ~PlanThreads() {
    if (outer) {
      tl_plan_threads = 0;
      --g_active_plans;
    }
  }
};

}  // namespace

static constexpr int64_t kMaxI32 = 2147483647LL;

bool csc_to_csr(int64_t n, const int64_t* colptr, const int64_t* rowval,
                const double* nzval, int base, HostCsr& out,
                std::string& err) {
  if (n < 1 || n >= kMaxI32) {
    err = "upload_csc: n out of range";
    return false;
  }
  if (base != 0 && base != 1) {
    err = "upload_csc: index_base must be 0 or 1";
    return false;
  }
  if (!colptr || !rowval || !nzval) {
    err = "upload_csc: null array";
    return false;
  }
  const int64_t nnz = colptr[n] - base;
  if (colptr[0] != base || nnz < 0 || nnz >= kMaxI32) {
    err = "upload_csc: malformed colptr or nnz >= 2^31";
    return false;
  }
  out.n = n;
  out.ptr.assign(n + 1, 0);
  for (int64_t j = 0; j < n; ++j) {
    if (colptr[j + 1] < colptr[j]) {
      err = "upload_csc: colptr not monotone";
      return false;
    }
  }
  // Symmetric: column j of the CSC is row j of the CSR.
  out.idx.resize(nnz);
  out.val.resize(nnz);
  for (int64_t j = 0; j <= n; ++j) out.ptr[j] = (int32_t)(colptr[j] - base);
  for (int64_t k = 0; k < nnz; ++k) {
    const int64_t r = rowval[k] - base;
    if (r < 0 || r >= n) {
      err = "upload_csc: row index out of range";
      return false;
    }
    out.idx[k] = (int32_t)r;
    out.val[k] = nzval[k];
  }
  return true;
}

namespace {

int64_t pow2_at_least(int64_t v) {
  int64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

int log2_exact(int64_t v) {
  int l = 0;
  while ((int64_t(1) << l) < v) ++l;
  return l;
}

// Lanes one row of dg entries takes at E = 2^elog entries per lane: chunks
// of at most 64 lanes (64 E entries), each in an aligned power-of-two
// segment.
int64_t row_lanes(int64_t dg, int elog) {
  if (dg == 0) return 0;
  const int64_t E = int64_t(1) << elog, per = 64 * E;
  const int64_t rem = dg % per;
  return 64 * (dg / per) + (rem ? pow2_at_least((rem + E - 1) / E) : 0);
}

// Entries per lane of a program from the mean length of its non-empty
// rows: more entries per lane for longer rows, so that a long row takes
// fewer lanes and fewer cross-lane reduction stages.
// Entries per lane from the mean row length.  Eight entries per lane
// (fewer segment sums for long rows) measured slower on MI355X: the tail of
// grid3(128) (81 entries per row) 0.94 -> 1.11 ms per iteration, that of
// path + 4 matchings n = 1e6 (495 per row) 5.73 -> 6.23 ms.
int choose_elog(int64_t rows, int64_t ent, int max_elog) {
  if (rows == 0) return 0;
  const double avg = (double)ent / (double)rows;
  int e = avg >= 16.0 ? 2 : avg >= 8.0 ? 1 : 0;
  return std::min(e, max_elog);
}

struct RowEntries {
  int32_t row;  // local
  std::vector<std::pair<int32_t, double>> e;  // (local src, weight)
};

// Entries [k0, k1) of one row, packed as chunks of at most 64 E entries.
// A piece of several chunks whose leads store partial sums (split) gets a
// fix-up row that adds them in chunk order.
struct Piece {
  const RowEntries* R;
  int64_t k0, k1;
  bool split;
};

// One pass: the pieces' chunks in steps of `lanes` lanes, the last step
// ending with a barrier.  Chunk leads add to y[row] (lead = 1) when the
// piece is one chunk, store to the partial slot part_base + i (lead = 3,
// i counting the split pieces' chunks in piece order) when it is split,
// and add atomically (lead = 2) for the chunks of an unsplit multi-chunk
// piece (the order-free form).  Lane j of a chunk takes its entries
// j E .. j E + E - 1.  Chunks are placed largest first with a bump
// pointer per step, which keeps every chunk aligned to its size.  The
// split pieces' fix-up rows are appended to *fix.
void pack_pass(const std::vector<Piece>& pieces, int elog, int64_t lanes,
               std::vector<TailStep>& steps, std::vector<TailSlot>& slots, int64_t part_base,
               std::vector<RowEntries>* fix) {
  const int64_t E = int64_t(1) << elog;
  struct Chunk {
    uint32_t dst;  // row, or partial slot
    uint32_t lead;
    int64_t size;  // segment lanes
    int64_t k0, k1;
    const RowEntries* R;
  };
  std::vector<Chunk> chunks;
  int64_t npart = 0;
  for (const Piece& Q : pieces) {
    const bool multi = Q.k1 - Q.k0 > 64 * E;
    RowEntries* F = nullptr;
    if (Q.split) {
      fix->push_back(RowEntries{Q.R->row, {}});
      F = &fix->back();
    }
    for (int64_t k0 = Q.k0; k0 < Q.k1; k0 += 64 * E) {
      const int64_t k1 = std::min(Q.k1, k0 + 64 * E);
      uint32_t dst = (uint32_t)Q.R->row, lead = multi ? 2u : 1u;
      if (Q.split) {
        dst = (uint32_t)(part_base + npart);
        lead = 3u;
        F->e.emplace_back((int32_t)(part_base + npart), 1.0);
        ++npart;
      }
      chunks.push_back({dst, lead, pow2_at_least((k1 - k0 + E - 1) / E), k0, k1, Q.R});
    }
  }
  std::stable_sort(chunks.begin(), chunks.end(),
                   [](const Chunk& a, const Chunk& b) { return a.size > b.size; });
  // placement: pass (step) p and lane offset of every chunk, bump pointers
  std::vector<int64_t> ptr, at(chunks.size());
  std::vector<int> maxlog;
  std::vector<int32_t> in(chunks.size());
  for (size_t c = 0; c < chunks.size(); ++c) {
    const Chunk& C = chunks[c];
    size_t p = 0;
    while (p < ptr.size() && ptr[p] + C.size > lanes) ++p;
    if (p == ptr.size()) {
      ptr.push_back(0);
      maxlog.push_back(0);
    }
    in[c] = (int32_t)p;
    at[c] = ptr[p];
    ptr[p] += C.size;
    maxlog[p] = std::max(maxlog[p], log2_exact(C.size));
  }
  // step p: E planes of cnt_p slots, plane-major, written in place (unused
  // slots stay {0, 0, 0})
  std::vector<int64_t> base(ptr.size()), cnt(ptr.size());
  for (size_t p = 0; p < ptr.size(); ++p) {
    cnt[p] = (ptr[p] + 63) / 64 * 64;
    base[p] = (int64_t)slots.size();
    steps.push_back(TailStep{(int32_t)base[p],
                             step_meta(cnt[p], p + 1 == ptr.size(), elog, maxlog[p])});
    slots.resize(slots.size() + E * cnt[p], TailSlot{0.0, 0, 0});
  }
  for (size_t c = 0; c < chunks.size(); ++c) {
    const Chunk& C = chunks[c];
    const int64_t p = in[c], off = at[c];
    const int lg = log2_exact(C.size);
    TailSlot* S0 = slots.data() + base[p];
    for (int64_t t = 0; t < C.size; ++t) {
      S0[off + t].meta = slot_meta(C.dst, (uint32_t)lg, t == C.size - 1 ? C.lead : 0u);
      for (int64_t e = 0; e < E; ++e) {
        const int64_t k = C.k0 + t * E + e;
        if (k < C.k1) {
          TailSlot& S = S0[e * cnt[p] + off + t];
          S.src = (uint32_t)C.R->e[k].first;
          S.w = C.R->e[k].second;
        }
      }
    }
  }
}

// Pack one level's rows (see TailSlot), E = 2^elog entries per lane; every
// pass ends with a barrier.  Without det, a row of at most 64 E entries is
// one chunk whose lead adds its sum to y[row], and a longer row's chunks
// add with atomics, in whatever order they land.  With det every sum is
// formed in a fixed order: a longer row's chunks store partial sums in
// slots part_base + i (i < kPartCap, past the program's own indices), and
// a fix-up pass (packed the same way, recursively) adds each row's
// partials in chunk order; rows whose chunks do not fit the slots of one
// pass go on in later passes (each pass adds to y[row] once).  Returns the
// partial slots used.
int64_t pack_level(std::vector<RowEntries>& rows, int elog, int64_t lanes,
                   std::vector<TailStep>& steps, std::vector<TailSlot>& slots, bool det,
                   int64_t part_base) {
  const int64_t E = int64_t(1) << elog, CH = 64 * E;
  if (!det) {
    std::vector<Piece> pieces;
    for (const RowEntries& R : rows)
      if (!R.e.empty()) pieces.push_back({&R, 0, (int64_t)R.e.size(), false});
    pack_pass(pieces, elog, lanes, steps, slots, 0, nullptr);
    return 0;
  }
  // chunks with partial slots per pass, leaving room for the fix-up
  // passes' own partials (a fix-up row longer than 64 E entries)
  const int64_t cap = kPartCap - kPartCap / 4;
  std::vector<std::vector<Piece>> passes(1);
  std::vector<int64_t> used(1, 0);
  for (const RowEntries& R : rows) {
    const int64_t dg = (int64_t)R.e.size();
    if (dg == 0) continue;
    if (dg <= CH) {
      passes[0].push_back({&R, 0, dg, false});
      continue;
    }
    size_t p = 0;  // first pass this row's next piece may use
    for (int64_t k0 = 0; k0 < dg;) {
      while (p < passes.size() && used[p] >= cap) ++p;
      if (p == passes.size()) {
        passes.emplace_back();
        used.push_back(0);
      }
      const int64_t k1 = std::min(dg, k0 + (cap - used[p]) * CH);
      const int64_t nch = (k1 - k0 + CH - 1) / CH;
      passes[p].push_back({&R, k0, k1, nch > 1});
      if (nch > 1) used[p] += nch;
      k0 = k1;
      ++p;
    }
  }
  int64_t npart = 0;
  for (size_t p = 0; p < passes.size(); ++p) {
    std::vector<RowEntries> fix;
    pack_pass(passes[p], elog, lanes, steps, slots, part_base, &fix);
    int64_t nested = 0;
    if (!fix.empty()) nested = pack_level(fix, elog, lanes, steps, slots, true, part_base + used[p]);
    npart = std::max(npart, used[p] + nested);
  }
  return npart;
}

}  // namespace

bool build_tail(Plan& P, int64_t max_steps, int64_t cap, int max_elog, int lanes,
                const DenseModel& dm, bool det) {
  const int64_t n = P.n;
  // extent: the longest suffix of levels within cap whose levels need at
  // most max_steps steps of kExtentLanes lanes at one entry per lane
  int64_t t = P.nlev, nt = 0;
  while (t > 0) {
    const int64_t l = t - 1;
    const int64_t w = P.lvl[l + 1] - P.lvl[l];
    if (nt + w > cap) break;
    int64_t fl = 0, bl = 0;
    for (int64_t pos = P.lvl[l]; pos < P.lvl[l + 1]; ++pos) {
      fl += row_lanes(P.Fptr[pos + 1] - P.Fptr[pos], 0);
      bl += row_lanes(P.Bptr[pos + 1] - P.Bptr[pos], 0);
    }
    // bump packing of power-of-two segments in decreasing size fills
    // whole steps, so lanes / kTailLanes rounded up is the step count
    const int64_t steps_l = (std::max(fl, bl) + kExtentLanes - 1) / kExtentLanes;
    if (steps_l > max_steps) break;
    nt += w;
    t = l;
  }
  TailProgram& T = P.tail;
  T = TailProgram();
  P.dense = DenseBlock();
  P.tail_level = t;
  T.T0 = P.lvl[t];
  T.NT = n - T.T0;
  T.NT1 = T.NT;
  const int32_t T0 = (int32_t)T.T0;
  T.Eptr.assign(T.NT + 1, 0);
  int64_t rows = 0, ent = 0;
  for (int64_t i = 0; i < T.NT; ++i) {
    const int64_t pos = T0 + i;
    int64_t fin = 0;
    for (int32_t k = P.Fptr[pos]; k < P.Fptr[pos + 1]; ++k)
      if (P.Fidx[k] < T0) {
        T.Eidx.push_back(P.Fidx[k]);
        T.Ew.push_back(P.Fw[k]);
      } else {
        ++fin;
      }
    T.Eptr[i + 1] = (int32_t)T.Eidx.size();
    const int64_t bin = P.Bptr[pos + 1] - P.Bptr[pos];
    rows += (fin > 0) + (bin > 0);
    ent += fin + bin;
  }
  T.elog = choose_elog(rows, ent, max_elog);
  struct LevelProg {
    std::vector<TailStep> steps;
    std::vector<TailSlot> slots;
    int64_t npart = 0;  // partial slots (pack_level)
  };
  PlanThreads threads;
  Team team(threads.n);
  // programs of levels [l0, nlev) with indices local to `base`: forward rows
  // gather their in-edges from [base, n), backward rows all their entries
  // (every level packs on its own, offsets local)
  std::atomic<bool> ok{true};
  // partial slots follow the program's own indices [0, n - base)
  auto pack_levels = [&](int64_t l0, int64_t base, int elog, std::vector<LevelProg>& fp,
                         std::vector<LevelProg>& bp) {
    const int64_t nl = P.nlev - l0;
    fp.assign(nl, LevelProg());
    bp.assign(nl, LevelProg());
    team.for_each(2 * nl, 4, [&](int64_t lo, int64_t hi) {
      std::vector<RowEntries> rws;
      for (int64_t q = lo; q < hi; ++q) {
        const bool fwd = q < nl;
        const int64_t l = l0 + (fwd ? q : q - nl);
        rws.clear();
        for (int64_t pos = P.lvl[l]; pos < P.lvl[l + 1]; ++pos) {
          RowEntries R;
          R.row = (int32_t)(pos - base);
          if (fwd) {
            for (int32_t k = P.Fptr[pos]; k < P.Fptr[pos + 1]; ++k)
              if (P.Fidx[k] >= base) R.e.emplace_back((int32_t)(P.Fidx[k] - base), P.Fw[k]);
          } else {
            for (int32_t k = P.Bptr[pos]; k < P.Bptr[pos + 1]; ++k) {
              if (P.Bidx[k] < base) ok = false;
              R.e.emplace_back((int32_t)(P.Bidx[k] - base), P.Bw[k]);
            }
          }
          rws.push_back(std::move(R));
        }
        LevelProg& G = fwd ? fp[l - l0] : bp[l - l0];
        G.npart = pack_level(rws, elog, lanes, G.steps, G.slots, det, n - base);
      }
    });
  };
  std::vector<LevelProg> fp, bp;
  pack_levels(t, T0, T.elog, fp, bp);
  const int64_t nl = P.nlev - t;

  // ---- the dense block: the split level by the cost model -----------------
  int64_t L1 = P.nlev;  // first level of the dense block (nlev: none)
  LevelProg bnd;        // the block's inputs from [T0, T1), one more level
  if (dm.max_k > 0 && nl > 0) {
    const double E = (double)(int64_t(1) << T.elog);
    std::vector<int64_t> steps(nl + 1, 0);  // steps of levels [t, t + q)
    for (int64_t q = 0; q < nl; ++q)
      steps[q + 1] = steps[q] + (int64_t)(fp[q].steps.size() + bp[q].steps.size());
    double best = (double)steps[nl] * dm.step_us;
    // entries from the sparse part [T0, T1) into the block, kept as the
    // split moves down one level at a time
    int64_t cross = 0;
    double smin = INFINITY, smax = 0.0;
    for (int64_t l = P.nlev - 1; l >= t; --l) {
      const int64_t lo = P.lvl[l], hi = P.lvl[l + 1];
      const int64_t K = n - lo;
      if (K > dm.max_k) break;
      for (int64_t pos = lo; pos < hi; ++pos) {
        const double sv = std::fabs(P.s[pos]);
        smin = std::min(smin, sv);
        smax = std::max(smax, sv);
      }
      if (!(smax <= dm.max_spread * smin)) break;
      // level l joins the block: its out-edges into the block stop crossing,
      // its in-edges from [T0, lo) start to
      for (int64_t pos = lo; pos < hi; ++pos) {
        for (int32_t k = P.Bptr[pos]; k < P.Bptr[pos + 1]; ++k)
          if (P.Bidx[k] >= hi) --cross;
        for (int32_t k = P.Fptr[pos]; k < P.Fptr[pos + 1]; ++k)
          if (P.Fidx[k] >= T0 && P.Fidx[k] < lo) ++cross;
      }
      const int64_t q = l - t;  // sparse levels [t, l)
      const double bsteps = std::ceil((double)cross / (E * (double)lanes * 0.7));
      const double cost = ((double)steps[q] + bsteps) * dm.step_us +
                          8.0 * (double)K * (double)K / (dm.gbps * 1e3) +
                          (q > 0 ? 2.0 * dm.launch_us : 0.0);
      if (cost < best) {
        best = cost;
        L1 = l;
      }
    }
  }
  const int64_t T1 = P.lvl[L1];
  T.NT1 = T1 - T0;
  if (L1 < P.nlev) {
    // the block's inputs from the sparse part, packed as one more level
    std::vector<RowEntries> rws;
    for (int64_t pos = T1; pos < n; ++pos) {
      RowEntries R;
      R.row = (int32_t)(pos - T0);
      for (int32_t k = P.Fptr[pos]; k < P.Fptr[pos + 1]; ++k)
        if (P.Fidx[k] >= T0 && P.Fidx[k] < T1) R.e.emplace_back((int32_t)(P.Fidx[k] - T0), P.Fw[k]);
      if (!R.e.empty()) rws.push_back(std::move(R));
    }
    if (!rws.empty()) bnd.npart = pack_level(rws, T.elog, lanes, bnd.steps, bnd.slots, det, T.NT);
    // the block's own programs, local to T1
    DenseBlock& D = P.dense;
    D.T1 = T1;
    D.K = n - T1;
    D.level = L1;
    int64_t drows = 0, dent = 0;
    for (int64_t pos = T1; pos < n; ++pos) {
      int64_t fin = 0;
      for (int32_t k = P.Fptr[pos]; k < P.Fptr[pos + 1]; ++k) fin += P.Fidx[k] >= T1;
      const int64_t bin = P.Bptr[pos + 1] - P.Bptr[pos];
      drows += (fin > 0) + (bin > 0);
      dent += fin + bin;
    }
    D.elog = choose_elog(drows, dent, max_elog);
    std::vector<LevelProg> dfp, dbp;
    pack_levels(L1, T1, D.elog, dfp, dbp);
    for (size_t q = 0; q < dfp.size(); ++q) D.npart = std::max({D.npart, dfp[q].npart, dbp[q].npart});
    for (size_t q = 0; q < dfp.size(); ++q) {
      const int32_t b0 = (int32_t)D.fslots.size();
      for (TailStep st : dfp[q].steps) {
        st.off += b0;
        D.fsteps.push_back(st);
      }
      D.fslots.insert(D.fslots.end(), dfp[q].slots.begin(), dfp[q].slots.end());
    }
    for (size_t q = dbp.size(); q-- > 0;) {
      const int32_t b0 = (int32_t)D.bslots.size();
      for (TailStep st : dbp[q].steps) {
        st.off += b0;
        D.bsteps.push_back(st);
      }
      D.bslots.insert(D.bslots.end(), dbp[q].slots.begin(), dbp[q].slots.end());
    }
  }
  // forward sweep: the sparse levels ascending (then the block's inputs);
  // backward: the sparse levels descending
  const int64_t ns = L1 - t;
  auto append = [](const LevelProg& G, std::vector<TailStep>& steps,
                   std::vector<TailSlot>& slots) {
    const int32_t base = (int32_t)slots.size();
    for (TailStep st : G.steps) {
      st.off += base;
      steps.push_back(st);
    }
    slots.insert(slots.end(), G.slots.begin(), G.slots.end());
  };
  size_t fs = bnd.steps.size(), fl = bnd.slots.size(), bs = 0, bl = 0;
  for (int64_t q = 0; q < ns; ++q) {
    fs += fp[q].steps.size();
    fl += fp[q].slots.size();
    bs += bp[q].steps.size();
    bl += bp[q].slots.size();
  }
  T.npart = bnd.npart;
  for (int64_t q = 0; q < ns; ++q) T.npart = std::max({T.npart, fp[q].npart, bp[q].npart});
  if (T.npart > kPartCap || P.dense.npart > kPartCap) return false;
  T.fsteps.reserve(fs);
  T.fslots.reserve(fl);
  T.bsteps.reserve(bs);
  T.bslots.reserve(bl);
  for (int64_t q = 0; q < ns; ++q) append(fp[q], T.fsteps, T.fslots);
  append(bnd, T.fsteps, T.fslots);
  for (int64_t q = ns - 1; q >= 0; --q) append(bp[q], T.bsteps, T.bslots);
  return ok.load();
}

namespace {

// One composite row: its terms in insertion order, merged by index in a
// per-thread table of positions (-1: index absent), so that equal indices
// are summed in insertion order starting from the first term; the row is
// emitted in index order.
std::vector<int32_t>& merge_where() {
  thread_local std::vector<int32_t> where;
  return where;
}

// Sort distinct values that arrive as ascending runs (a composite row is a
// union of sorted rows): pairwise merges of neighbouring runs, O(m log runs).
void sort_runs(int32_t* a, size_t m, std::vector<int32_t>& tmp, std::vector<size_t>& bounds) {
  bounds.clear();
  bounds.push_back(0);
  for (size_t k = 1; k < m; ++k)
    if (a[k] < a[k - 1]) bounds.push_back(k);
  bounds.push_back(m);
  if (bounds.size() <= 2) return;
  tmp.resize(m);
  int32_t *src = a, *dst = tmp.data();
  while (bounds.size() > 2) {
    size_t nb = 0, r = 0;
    for (; r + 2 < bounds.size(); r += 2) {
      std::merge(src + bounds[r], src + bounds[r + 1], src + bounds[r + 1], src + bounds[r + 2],
                 dst + bounds[r]);
      bounds[nb++] = bounds[r];
    }
    if (r + 1 < bounds.size()) {
      std::copy(src + bounds[r], src + bounds[r + 1], dst + bounds[r]);
      bounds[nb++] = bounds[r];
    }
    bounds[nb++] = m;
    bounds.resize(nb);
    std::swap(src, dst);
  }
  if (src != a) std::copy(src, src + m, a);
}

struct RowMerge {
  std::vector<int32_t>& where;
  std::vector<int32_t> ti, tmp;
  std::vector<size_t> bounds;
  std::vector<double> tw;
  explicit RowMerge(int64_t span) : where(merge_where()) {
    if ((int64_t)where.size() < span) where.assign((size_t)span, -1);
  }
  void begin() {
    ti.clear();
    tw.clear();
  }
  void add(int32_t i, double w) {
    int32_t& p = where[i];
    if (p < 0) {
      p = (int32_t)ti.size();
      ti.push_back(i);
      tw.push_back(w);
    } else {
      tw[p] += w;
    }
  }
  // append the merged row; returns its length (the table is left all -1)
  int32_t take(std::vector<int32_t>& idx, std::vector<double>& w) {
    const size_t m = ti.size(), at = idx.size();
    idx.insert(idx.end(), ti.begin(), ti.end());
    sort_runs(idx.data() + at, m, tmp, bounds);
    w.resize(at + m);
    for (size_t k = 0; k < m; ++k) {
      const int32_t i = idx[at + k];
      w[at + k] = tw[where[i]];
      where[i] = -1;
    }
    return (int32_t)m;
  }
};

// Rows of consecutive positions computed by one task.
struct RowChunk {
  std::vector<int32_t> idx, len;
  std::vector<double> w;
  int64_t orig = 0;
};

// The composite rows of one sweep while they are built: the chunks stay in
// place, and each position points at its row inside one of them.
struct RowTable {
  std::vector<std::unique_ptr<RowChunk>> chunks;
  std::vector<const int32_t*> idx;
  std::vector<const double*> w;
  std::vector<int32_t> len;
  explicit RowTable(int64_t rows) : idx(rows, nullptr), w(rows, nullptr), len(rows, 0) {}
};

}  // namespace

void build_blocks(Plan& P, int max_depth, double growth, int64_t max_extra) {
  const int64_t n = P.n;
  const int64_t L = P.tail_level;
  LevelBlocks& B = P.blocks;
  B = LevelBlocks();
  const int64_t T0 = P.lvl[L];
  B.fptr.assign(T0 + 1, 0);
  B.bptr.assign(T0 + 1, 0);
  if (L == 0) {
    B.lev = {0};
    return;
  }
  PlanThreads threads;
  Team team(threads.n);
  // rows [lo, hi) in chunks of `per` rows: row(v, rm) adds v's terms and
  // returns its original entries
  using Chunks = std::vector<std::unique_ptr<RowChunk>>;
  auto run_rows = [&](int64_t lo, int64_t hi, int64_t per, Chunks& out, const auto& row) {
    const int64_t nc = (hi - lo + per - 1) / per;
    out.resize(nc);
    team.for_each(nc, 1, [&](int64_t c0, int64_t c1) {
      RowMerge rm(n + T0);
      for (int64_t c = c0; c < c1; ++c) {
        if (!out[c]) out[c] = std::make_unique<RowChunk>();
        RowChunk& O = *out[c];
        O.idx.clear();
        O.w.clear();
        O.len.clear();
        O.orig = 0;
        const int64_t a = lo + c * per, b = std::min(hi, a + per);
        for (int64_t v = a; v < b; ++v) {
          rm.begin();
          O.orig += row(v, rm);
          O.len.push_back(rm.take(O.idx, O.w));
        }
      }
    });
  };
  // rows per chunk: about four chunks per thread, 16..256 rows
  auto per_chunk = [&](int64_t rows) {
    const int64_t want = (rows + 4 * team.size() - 1) / (4 * team.size());
    return std::max<int64_t>(16, std::min<int64_t>(256, want));
  };
  // hand the chunks of rows [lo, ...) to the table
  auto commit = [](RowTable& T, int64_t lo, int64_t per, Chunks& cs) {
    for (size_t c = 0; c < cs.size(); ++c) {
      const RowChunk& O = *cs[c];
      int64_t v = lo + (int64_t)c * per, e = 0;
      for (int32_t len : O.len) {
        T.idx[v] = O.idx.data() + e;
        T.w[v] = O.w.data() + e;
        T.len[v] = len;
        e += len;
        ++v;
      }
      T.chunks.push_back(std::move(cs[c]));
    }
    cs.clear();
  };
  // the table's rows [0, T0) in position order
  auto copy_out = [&](const RowTable& T, std::vector<int32_t>& ptr, std::vector<int32_t>& idx,
                      std::vector<double>& w) {
    for (int64_t v = 0; v < T0; ++v) ptr[v + 1] = ptr[v] + T.len[v];
    idx.resize((size_t)ptr[T0]);
    w.resize((size_t)ptr[T0]);
    team.for_each(T0, 4096, [&](int64_t a, int64_t b) {
      for (int64_t v = a; v < b; ++v) {
        std::copy(T.idx[v], T.idx[v] + T.len[v], idx.begin() + ptr[v]);
        std::copy(T.w[v], T.w[v] + T.len[v], w.begin() + ptr[v]);
      }
    });
  };

  // forward composites and the grouping (forward rows of a group's earlier
  // levels do not change when the group grows at its end)
  RowTable F(T0);
  Chunks cur;
  int64_t L0 = 0;
  B.lev.push_back(0);
  int64_t orig_block = 0, comp_block = 0;
  for (int64_t l = 0; l < L; ++l) {
    const int64_t lo = P.lvl[l], hi = P.lvl[l + 1];
    const int64_t per = per_chunk(hi - lo);
    int64_t orig = 0, comp = 0;
    auto compose = [&](int64_t start) {
      const int32_t s0 = P.lvl[start];
      run_rows(lo, hi, per, cur, [&](int64_t v, RowMerge& rm) -> int64_t {
        rm.add((int32_t)(n + v), 1.0);
        for (int32_t k = P.Fptr[v]; k < P.Fptr[v + 1]; ++k) {
          const int32_t c = P.Fidx[k];
          if (c < s0) {
            rm.add(c, P.Fw[k]);
          } else {
            const int32_t* ci = F.idx[c];
            const double* cw = F.w[c];
            for (int32_t e = 0; e < F.len[c]; ++e) rm.add(ci[e], P.Fw[k] * cw[e]);
          }
        }
        return P.Fptr[v + 1] - P.Fptr[v] + 1;
      });
      orig = comp = 0;
      for (const auto& O : cur) {
        orig += O->orig;
        comp += (int64_t)O->idx.size();
      }
    };
    bool joined = false;
    if (l > L0 && l - L0 < max_depth) {
      compose(L0);
      if (comp <= growth * (double)orig && comp - orig <= max_extra &&
          comp_block + comp <= growth * (double)(orig_block + orig)) {
        joined = true;
        orig_block += orig;
        comp_block += comp;
      }
    }
    if (!joined) {
      L0 = l;
      if (l > 0) B.lev.push_back((int32_t)l);
      compose(L0);  // a fresh group: every in-edge is outside it
      orig_block = orig;
      comp_block = comp;
    }
    commit(F, lo, per, cur);
  }
  B.lev.push_back((int32_t)L);
  copy_out(F, B.fptr, B.fidx, B.fw);
  F = RowTable(0);
  // backward composites per group, top level first: a row's in-group
  // terms expand into the rows of the group's higher levels
  RowTable Bk(T0);
  for (size_t g = 0; g + 1 < B.lev.size(); ++g) {
    const int64_t g0 = B.lev[g], g1 = B.lev[g + 1];
    const int32_t top = P.lvl[g1];
    for (int64_t l = g1 - 1; l >= g0; --l) {
      const int64_t lo = P.lvl[l], hi = P.lvl[l + 1];
      const int64_t per = per_chunk(hi - lo);
      run_rows(lo, hi, per, cur, [&](int64_t v, RowMerge& rm) -> int64_t {
        rm.add((int32_t)(n + v), P.s[v]);
        for (int32_t k = P.Bptr[v]; k < P.Bptr[v + 1]; ++k) {
          const int32_t j = P.Bidx[k];
          if (j >= top) {
            rm.add(j, P.Bw[k]);
          } else {
            const int32_t* ji = Bk.idx[j];
            const double* jw = Bk.w[j];
            for (int32_t e = 0; e < Bk.len[j]; ++e) rm.add(ji[e], P.Bw[k] * jw[e]);
          }
        }
        return 0;
      });
      commit(Bk, lo, per, cur);
    }
  }
  copy_out(Bk, B.bptr, B.bidx, B.bw);
  merge_where() = std::vector<int32_t>();
}

bool build_plan(const HostCsr& L, int64_t nnzF, const int64_t* col,
                const int64_t* colptr, const int64_t* rowval,
                const double* fval, const double* d, int base,
                int64_t max_tail_steps, Plan& P, std::string& err,
                const DenseModel& dense) {
  const int64_t n = L.n;
  if (n < 1) {
    err = "upload_ldli: upload the Laplacian first";
    return false;
  }
  if (nnzF < 0 || nnzF >= kMaxI32) {
    err = "upload_ldli: nnzF out of range";
    return false;
  }
  if (n > 1 && (!col || !colptr || !rowval || !fval || !d)) {
    err = "upload_ldli: null array";
    return false;
  }
  // ---- validate the factor and find the elimination position -------------
  std::vector<int32_t> elim(n, -1);
  for (int64_t ii = 0; ii < n - 1; ++ii) {
    const int64_t v = col[ii] - base;
    if (v < 0 || v >= n || elim[v] != -1) {
      err = "upload_ldli: col is not a list of distinct vertices";
      return false;
    }
    elim[v] = (int32_t)ii;
  }
  int64_t root = -1;
  for (int64_t v = 0; v < n; ++v)
    if (elim[v] < 0) {
      root = v;
      elim[v] = (int32_t)(n - 1);
    }
  PlanThreads threads;
  Team team(threads.n);
  constexpr int64_t kGrain = 4096;
  if (n > 1) {
    if (colptr[0] != base || colptr[n - 1] - base != nnzF) {
      err = "upload_ldli: colptr does not span nnzF entries";
      return false;
    }
    // the first bad column (in column order) names the error, as a
    // sequential scan would
    int64_t bad_col = INT64_MAX;
    int bad_kind = 0;
    std::mutex bad_mu;
    team.for_each(n - 1, kGrain, [&](int64_t lo, int64_t hi) {
      for (int64_t ii = lo; ii < hi; ++ii) {
        int kind = 0;
        if (colptr[ii + 1] <= colptr[ii]) {
          kind = 1;
        } else {
          for (int64_t k = colptr[ii] - base; k < colptr[ii + 1] - base; ++k) {
            const int64_t j = rowval[k] - base;
            if (j < 0 || j >= n || elim[j] <= ii) {
              kind = 2;
              break;
            }
          }
        }
        if (kind) {
          std::lock_guard<std::mutex> g(bad_mu);
          if (ii < bad_col) {
            bad_col = ii;
            bad_kind = kind;
          }
          return;
        }
      }
    });
    if (bad_kind == 1) {
      err = "upload_ldli: empty or non-monotone column";
      return false;
    }
    if (bad_kind == 2) {
      err = "upload_ldli: a row of a column is not eliminated later";
      return false;
    }
  }
  (void)root;
  P.n = n;
  P.nnzF = nnzF;

  // ---- per-entry weights and per-column scale -----------------------------
  std::vector<double> w(nnzF);
  std::vector<double> Pc(n, 1.0);
  team.for_each(n - 1, kGrain, [&](int64_t lo, int64_t hi) {
    for (int64_t ii = lo; ii < hi; ++ii) {
      const int64_t k0 = colptr[ii] - base, k1 = colptr[ii + 1] - base - 1;
      double prod = 1.0;
      for (int64_t k = k0; k < k1; ++k) {
        w[k] = fval[k] * prod;
        prod *= (1.0 - fval[k]);
      }
      w[k1] = prod;
      Pc[col[ii] - base] = prod;
    }
  });

  // ---- forward levels (longest path from a source) ------------------------
  std::vector<int32_t> lev(n, 0);
  for (int64_t ii = 0; ii < n - 1; ++ii) {
    const int64_t v = col[ii] - base;
    const int32_t l = lev[v] + 1;
    for (int64_t k = colptr[ii] - base; k < colptr[ii + 1] - base; ++k) {
      const int64_t j = rowval[k] - base;
      if (lev[j] < l) lev[j] = l;
    }
  }
  int32_t maxlev = 0;
  for (int64_t v = 0; v < n; ++v) maxlev = std::max(maxlev, lev[v]);
  P.nlev = (int64_t)maxlev + 1;
  P.lvl.assign(P.nlev + 1, 0);
  for (int64_t v = 0; v < n; ++v) P.lvl[lev[v] + 1]++;
  for (int64_t l = 0; l < P.nlev; ++l) P.lvl[l + 1] += P.lvl[l];

  P.perm.assign(n, 0);
  P.iperm.assign(n, 0);
  {
    std::vector<int32_t> cur(P.lvl.begin(), P.lvl.end() - 1);
    for (int64_t v = 0; v < n; ++v) {
      const int32_t pos = cur[lev[v]]++;
      P.perm[pos] = (int32_t)v;
      P.iperm[v] = pos;
    }
  }


  // ---- permuted Laplacian --------------------------------------------------
  const int64_t nnzL = (int64_t)L.idx.size();
  P.Lptr.assign(n + 1, 0);
  P.Lidx.resize(nnzL);
  P.Lval.resize(nnzL);
  for (int64_t pos = 0; pos < n; ++pos) {
    const int32_t v = P.perm[pos];
    P.Lptr[pos + 1] = P.Lptr[pos] + (L.ptr[v + 1] - L.ptr[v]);
  }
  team.for_each(n, kGrain, [&](int64_t lo, int64_t hi) {
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t pos = lo; pos < hi; ++pos) {
      const int32_t v = P.perm[pos];
      row.clear();
      for (int32_t k = L.ptr[v]; k < L.ptr[v + 1]; ++k)
        row.emplace_back(P.iperm[L.idx[k]], L.val[k]);
      std::sort(row.begin(), row.end(),
                [](const auto& a, const auto& b) { return a.first < b.first; });
      int32_t o = P.Lptr[pos];
      for (auto& e : row) {
        P.Lidx[o] = e.first;
        P.Lval[o] = e.second;
        ++o;
      }
    }
  });

  // ---- backward rows: column of v, in column order ------------------------
  P.Bptr.assign(n + 1, 0);
  P.Bidx.resize(nnzF);
  P.Bw.resize(nnzF);
  P.s.assign(n, 1.0);
  for (int64_t ii = 0; ii < n - 1; ++ii) {
    const int64_t v = col[ii] - base;
    P.Bptr[P.iperm[v] + 1] = (int32_t)(colptr[ii + 1] - colptr[ii]);
  }
  for (int64_t pos = 0; pos < n; ++pos) P.Bptr[pos + 1] += P.Bptr[pos];
  // destination row (position) of every factor entry, for both directions
  std::vector<int32_t> dest(nnzF);
  team.for_each(n - 1, kGrain, [&](int64_t lo, int64_t hi) {
    for (int64_t ii = lo; ii < hi; ++ii) {
      const int64_t v = col[ii] - base;
      int32_t o = P.Bptr[P.iperm[v]];
      for (int64_t k = colptr[ii] - base; k < colptr[ii + 1] - base; ++k) {
        const int32_t j = P.iperm[rowval[k] - base];
        dest[k] = j;
        P.Bidx[o] = j;
        P.Bw[o] = w[k];
        ++o;
      }
      const double dv = d[v];
      P.s[P.iperm[v]] = dv != 0.0 ? Pc[v] * Pc[v] / dv : Pc[v] * Pc[v];
    }
  });
  // the never-eliminated root: LDLsolver divides y[root] by d[root] when it
  // is nonzero (src/approxChol.jl:1351-1355); the reference's builders
  // leave it 0, a caller's LDLinv may not
  if (root >= 0 && n > 1 && d[root] != 0.0) P.s[P.iperm[root]] = 1.0 / d[root];

  // ---- forward rows: in-edges of v, in elimination order of the source ----
  P.Fptr.assign(n + 1, 0);
  P.Fidx.resize(nnzF);
  P.Fw.resize(nnzF);
  // a transpose: the columns are split into contiguous ranges, one per
  // part; each part counts its entries per row, the counts give every
  // (part, row) its own cursor, and each part fills its columns in order,
  // so a row holds its entries in column (elimination) order
  {
    const int64_t parts = std::min<int64_t>(team.size(), 8);
    std::vector<int32_t> cur((size_t)parts * n, 0);
    auto range = [&](int64_t q) { return (n - 1) * q / parts; };
    team.for_each(parts, 1, [&](int64_t q0, int64_t q1) {
      for (int64_t q = q0; q < q1; ++q) {
        int32_t* c = cur.data() + q * n;
        for (int64_t k = colptr[range(q)] - base; k < colptr[range(q + 1)] - base; ++k) ++c[dest[k]];
      }
    });
    for (int64_t pos = 0; pos < n; ++pos) {
      int32_t tot = 0;
      for (int64_t q = 0; q < parts; ++q) tot += cur[q * n + pos];
      P.Fptr[pos + 1] = P.Fptr[pos] + tot;
    }
    team.for_each(n, kGrain, [&](int64_t lo, int64_t hi) {
      for (int64_t pos = lo; pos < hi; ++pos) {
        int32_t at = P.Fptr[pos];
        for (int64_t q = 0; q < parts; ++q) {
          const int32_t c = cur[q * n + pos];
          cur[q * n + pos] = at;
          at += c;
        }
      }
    });
    team.for_each(parts, 1, [&](int64_t q0, int64_t q1) {
      for (int64_t q = q0; q < q1; ++q) {
        int32_t* c = cur.data() + q * n;
        for (int64_t ii = range(q); ii < range(q + 1); ++ii) {
          const int32_t src = P.iperm[col[ii] - base];
          for (int64_t k = colptr[ii] - base; k < colptr[ii + 1] - base; ++k) {
            const int32_t o = c[dest[k]]++;
            P.Fidx[o] = src;
            P.Fw[o] = w[k];
          }
        }
      }
    });
  }
  if (!build_tail(P, max_tail_steps, kTailCap, kTailMaxElog, kTailLanes, dense)) {
    err = "upload_ldli: internal error packing the tail program";
    return false;
  }
  return true;
}

}  // namespace lapgpu

namespace lapgpu {
std::recursive_mutex& device_lock() {
  static std::recursive_mutex m;
  return m;
}
}  // namespace lapgpu
