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
#include <cstring>
#include <numeric>

namespace lapgpu {

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

// Pack one level's rows into steps of `lanes` lanes (see TailSlot), E =
// 2^elog entries per lane.  A row of at most 64 E entries is one segment
// whose lead lane adds its sum to y[row] (lead = 1); a longer row is split
// into chunks whose leads add with an LDS atomic (lead = 2).  Lane j of a
// segment takes the chunk's entries j E .. j E + E - 1.  Segments are
// placed largest first with a bump pointer per step, which keeps every
// segment aligned to its size.
void pack_level(std::vector<RowEntries>& rows, int elog, int64_t lanes,
                std::vector<TailStep>& steps, std::vector<TailSlot>& slots) {
  const int64_t E = int64_t(1) << elog;
  struct Chunk {
    int32_t row;
    int64_t size;  // segment lanes
    int64_t k0, k1;
    const RowEntries* R;
    bool atomic;
  };
  std::vector<Chunk> chunks;
  for (const RowEntries& R : rows) {
    const int64_t dg = (int64_t)R.e.size();
    if (dg == 0) continue;
    const bool multi = dg > 64 * E;
    for (int64_t k0 = 0; k0 < dg; k0 += 64 * E) {
      const int64_t k1 = std::min(dg, k0 + 64 * E);
      chunks.push_back({R.row, pow2_at_least((k1 - k0 + E - 1) / E), k0, k1, &R, multi});
    }
  }
  std::stable_sort(chunks.begin(), chunks.end(),
                   [](const Chunk& a, const Chunk& b) { return a.size > b.size; });
  // pass p: E planes of kTailLanes slots, plane-major
  std::vector<std::vector<TailSlot>> pass;
  std::vector<int64_t> ptr;
  std::vector<int> maxlog;
  for (const Chunk& C : chunks) {
    size_t p = 0;
    while (p < pass.size() && ptr[p] + C.size > lanes) ++p;
    if (p == pass.size()) {
      pass.emplace_back(E * lanes, TailSlot{0.0, 0, 0});
      ptr.push_back(0);
      maxlog.push_back(0);
    }
    const int64_t off = ptr[p];
    ptr[p] += C.size;
    const int lg = log2_exact(C.size);
    maxlog[p] = std::max(maxlog[p], lg);
    for (int64_t t = 0; t < C.size; ++t) {
      TailSlot& M = pass[p][off + t];
      M.meta = slot_meta((uint32_t)C.row, (uint32_t)lg,
                         t == C.size - 1 ? (C.atomic ? 2u : 1u) : 0u);
      for (int64_t e = 0; e < E; ++e) {
        const int64_t k = C.k0 + t * E + e;
        if (k < C.k1) {
          TailSlot& S = pass[p][e * lanes + off + t];
          S.src = (uint32_t)C.R->e[k].first;
          S.w = C.R->e[k].second;
        }
      }
    }
  }
  for (size_t p = 0; p < pass.size(); ++p) {
    const int64_t cnt = (ptr[p] + 63) / 64 * 64;
    steps.push_back(TailStep{(int32_t)slots.size(),
                             step_meta(cnt, p + 1 == pass.size(), elog, maxlog[p])});
    for (int64_t e = 0; e < E; ++e)
      slots.insert(slots.end(), pass[p].begin() + e * lanes,
                   pass[p].begin() + e * lanes + cnt);
  }
}

}  // namespace

bool build_tail(Plan& P, int64_t max_steps, int64_t cap, int max_elog, int lanes) {
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
  P.tail_level = t;
  T.T0 = P.lvl[t];
  T.NT = n - T.T0;
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
  bool ok = true;
  for (int64_t l = t; l < P.nlev; ++l) {
    std::vector<RowEntries> rws;
    for (int64_t pos = P.lvl[l]; pos < P.lvl[l + 1]; ++pos) {
      RowEntries R;
      R.row = (int32_t)(pos - T0);
      for (int32_t k = P.Fptr[pos]; k < P.Fptr[pos + 1]; ++k)
        if (P.Fidx[k] >= T0) R.e.emplace_back(P.Fidx[k] - T0, P.Fw[k]);
      rws.push_back(std::move(R));
    }
    pack_level(rws, T.elog, lanes, T.fsteps, T.fslots);
  }
  for (int64_t l = P.nlev - 1; l >= t && ok; --l) {
    std::vector<RowEntries> rws;
    for (int64_t pos = P.lvl[l]; pos < P.lvl[l + 1]; ++pos) {
      RowEntries R;
      R.row = (int32_t)(pos - T0);
      for (int32_t k = P.Bptr[pos]; k < P.Bptr[pos + 1]; ++k) {
        if (P.Bidx[k] < T0) ok = false;
        R.e.emplace_back(P.Bidx[k] - T0, P.Bw[k]);
      }
      rws.push_back(std::move(R));
    }
    pack_level(rws, T.elog, lanes, T.bsteps, T.bslots);
  }
  return ok;
}

namespace {

// Sparse accumulator over [0, 2n) for composite rows.
struct Accum {
  std::vector<double> val;
  std::vector<int32_t> mark, touched;
  int32_t stamp = 0;
  explicit Accum(int64_t size) : val(size, 0.0), mark(size, -1) {}
  void begin() {
    ++stamp;
    touched.clear();
  }
  void add(int32_t i, double w) {
    if (mark[i] != stamp) {
      mark[i] = stamp;
      val[i] = 0.0;
      touched.push_back(i);
    }
    val[i] += w;
  }
};

using Rep = std::vector<std::pair<int32_t, double>>;

void take(Accum& A, Rep& out) {
  std::sort(A.touched.begin(), A.touched.end());
  out.clear();
  out.reserve(A.touched.size());
  for (int32_t i : A.touched) out.emplace_back(i, A.val[i]);
}

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
  Accum acc(2 * n);
  std::vector<Rep> frep(T0), brep(T0);
  // forward composites and the grouping (forward reps of a group's earlier
  // levels do not change when the group grows at its end)
  int64_t L0 = 0;
  B.lev.push_back(0);
  int64_t orig_block = 0, comp_block = 0;
  for (int64_t l = 0; l < L; ++l) {
    const int64_t lo = P.lvl[l], hi = P.lvl[l + 1];
    int64_t orig = 0, comp = 0;
    auto compose = [&](int64_t start) {
      orig = comp = 0;
      for (int64_t v = lo; v < hi; ++v) {
        acc.begin();
        acc.add((int32_t)(n + v), 1.0);
        for (int32_t k = P.Fptr[v]; k < P.Fptr[v + 1]; ++k) {
          const int32_t c = P.Fidx[k];
          if (c < P.lvl[start]) {
            acc.add(c, P.Fw[k]);
          } else {
            for (const auto& e : frep[c]) acc.add(e.first, P.Fw[k] * e.second);
          }
        }
        take(acc, frep[v]);
        orig += P.Fptr[v + 1] - P.Fptr[v] + 1;
        comp += (int64_t)frep[v].size();
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
  }
  B.lev.push_back((int32_t)L);
  // backward composites per group (top level first)
  for (size_t g = 0; g + 1 < B.lev.size(); ++g) {
    const int64_t g0 = B.lev[g], g1 = B.lev[g + 1];
    const int64_t top = P.lvl[g1];
    for (int64_t l = g1 - 1; l >= g0; --l) {
      for (int64_t v = P.lvl[l]; v < P.lvl[l + 1]; ++v) {
        acc.begin();
        acc.add((int32_t)(n + v), P.s[v]);
        for (int32_t k = P.Bptr[v]; k < P.Bptr[v + 1]; ++k) {
          const int32_t j = P.Bidx[k];
          if (j >= top) {
            acc.add(j, P.Bw[k]);
          } else {
            for (const auto& e : brep[j]) acc.add(e.first, P.Bw[k] * e.second);
          }
        }
        take(acc, brep[v]);
      }
    }
  }
  for (int64_t v = 0; v < T0; ++v) {
    B.fptr[v + 1] = B.fptr[v] + (int32_t)frep[v].size();
    B.bptr[v + 1] = B.bptr[v] + (int32_t)brep[v].size();
  }
  B.fidx.reserve(B.fptr[T0]);
  B.fw.reserve(B.fptr[T0]);
  B.bidx.reserve(B.bptr[T0]);
  B.bw.reserve(B.bptr[T0]);
  for (int64_t v = 0; v < T0; ++v) {
    for (const auto& e : frep[v]) {
      B.fidx.push_back(e.first);
      B.fw.push_back(e.second);
    }
    for (const auto& e : brep[v]) {
      B.bidx.push_back(e.first);
      B.bw.push_back(e.second);
    }
  }
}

bool build_plan(const HostCsr& L, int64_t nnzF, const int64_t* col,
                const int64_t* colptr, const int64_t* rowval,
                const double* fval, const double* d, int base,
                int64_t max_tail_steps, Plan& P, std::string& err) {
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
  if (n > 1) {
    if (colptr[0] != base || colptr[n - 1] - base != nnzF) {
      err = "upload_ldli: colptr does not span nnzF entries";
      return false;
    }
    for (int64_t ii = 0; ii < n - 1; ++ii) {
      if (colptr[ii + 1] <= colptr[ii]) {
        err = "upload_ldli: empty or non-monotone column";
        return false;
      }
      for (int64_t k = colptr[ii] - base; k < colptr[ii + 1] - base; ++k) {
        const int64_t j = rowval[k] - base;
        if (j < 0 || j >= n || elim[j] <= ii) {
          err = "upload_ldli: a row of a column is not eliminated later";
          return false;
        }
      }
    }
  }
  (void)root;
  P.n = n;
  P.nnzF = nnzF;

  // ---- per-entry weights and per-column scale -----------------------------
  std::vector<double> w(nnzF);
  std::vector<double> Pc(n, 1.0);
  for (int64_t ii = 0; ii < n - 1; ++ii) {
    const int64_t k0 = colptr[ii] - base, k1 = colptr[ii + 1] - base - 1;
    double prod = 1.0;
    for (int64_t k = k0; k < k1; ++k) {
      w[k] = fval[k] * prod;
      prod *= (1.0 - fval[k]);
    }
    w[k1] = prod;
    Pc[col[ii] - base] = prod;
  }

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
  {
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t pos = 0; pos < n; ++pos) {
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
  }

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
  for (int64_t ii = 0; ii < n - 1; ++ii) {
    const int64_t v = col[ii] - base;
    int32_t o = P.Bptr[P.iperm[v]];
    for (int64_t k = colptr[ii] - base; k < colptr[ii + 1] - base; ++k) {
      P.Bidx[o] = P.iperm[rowval[k] - base];
      P.Bw[o] = w[k];
      ++o;
    }
    const double dv = d[v];
    P.s[P.iperm[v]] = dv != 0.0 ? Pc[v] * Pc[v] / dv : Pc[v] * Pc[v];
  }

  // ---- forward rows: in-edges of v, in elimination order of the source ----
  P.Fptr.assign(n + 1, 0);
  P.Fidx.resize(nnzF);
  P.Fw.resize(nnzF);
  for (int64_t k = 0; k < nnzF; ++k) P.Fptr[P.iperm[rowval[k] - base] + 1]++;
  for (int64_t pos = 0; pos < n; ++pos) P.Fptr[pos + 1] += P.Fptr[pos];
  {
    std::vector<int32_t> cur(P.Fptr.begin(), P.Fptr.end() - 1);
    for (int64_t ii = 0; ii < n - 1; ++ii) {
      const int32_t src = P.iperm[col[ii] - base];
      for (int64_t k = colptr[ii] - base; k < colptr[ii + 1] - base; ++k) {
        const int32_t o = cur[P.iperm[rowval[k] - base]]++;
        P.Fidx[o] = src;
        P.Fw[o] = w[k];
      }
    }
  }
  if (!build_tail(P, max_tail_steps, kTailCap)) {
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
