// lapfactor.cpp — host restatement of the reference's randomized edge-by-edge
// elimination that produces LDLinv (src/approxChol.jl:864-977, greedy
// :deg order, split = merge = 0).  It is the stand-in for the reference's
// Julia factor build where no Julia is present; the GPU solve consumes its
// output exactly as it would consume the reference's LDLinv.
//
// Data structures (own layout, same semantics):
//  * every directed adjacency entry is one `Edge` in a flat array; each
//    vertex owns a singly linked list of edges whose tail links to itself
//    (src/approxChol.jl:102-136), and `rev` points at the twin (j,i) entry
//    (src/graphUtils.jl:49-55);
//  * the approximate min-degree queue is an array of doubly linked buckets
//    keyed by the multi-edge degree, with keys above n bundled as n + key/n
//    (src/approxChol.jl:2085-2239).
// Ties when sorting a column by weight are broken by ascending row, which is
// what the reference's insertion sort (used by Julia's QuickSort below 21
// elements) yields after its sort by row.

#include "lapfactor.h"

#include <algorithm>
#include <utility>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const char* msg) {
  g_err = msg;
  return code;
}

// xoshiro256++ seeded by splitmix64; uniform double in [0,1) from the top 53
// bits.  This is the build's own stream; it is not Julia's task-local RNG.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9e3779b97f4a7c15ULL;
      uint64_t t = z;
      t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ULL;
      t = (t ^ (t >> 27)) * 0x94d049bb133111ebULL;
      s[i] = t ^ (t >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// Stable sort of a short index list by key(x): insertion sort (no buffer
// allocation), std::stable_sort for long lists.  Both keep equal keys in
// their input order, so the result is the same.
template <class Key>
void stable_sort_by(std::vector<int64_t>& v, size_t len, const Key& key) {
  if (len > 48) {
    std::stable_sort(v.begin(), v.begin() + len,
                     [&](int64_t x, int64_t y) { return key(x) < key(y); });
    return;
  }
  for (size_t a = 1; a < len; ++a) {
    const int64_t x = v[a];
    const auto kx = key(x);
    size_t b = a;
    while (b > 0 && kx < key(v[b - 1])) {
      v[b] = v[b - 1];
      --b;
    }
    v[b] = x;
  }
}

struct Edge {
  int64_t row;
  double val;
  int64_t next;  // next edge in the owning vertex's list; == self at the tail
  int64_t rev;   // twin entry in the row vertex's list
};

// Bucketed approximate priority queue on integer keys.  Keys up to K are
// exact buckets, larger keys share bucket K + key / K (at most 2K + 1); K is
// n, or split * n for the split-edge elimination (keyMap,
// src/approxChol.jl:2085-2092, 2200-2239).
struct DegreeQueue {
  int64_t K;
  std::vector<int64_t> prev, next, key;  // per vertex; -1 = none
  std::vector<int64_t> head;             // per bucket
  int64_t minlist = 1;
  int64_t nitems;

  int64_t bucket(int64_t k) const {
    if (k <= K) return k;
    int64_t b = K + k / K;
    return b < 2 * K + 1 ? b : 2 * K + 1;
  }

  explicit DegreeQueue(const std::vector<int64_t>& deg, int64_t split = 1)
      : K((int64_t)deg.size() * split),
        prev(deg.size(), -1),
        next(deg.size(), -1),
        key(deg),
        head(2 * deg.size() * split + 2, -1),
        nitems((int64_t)deg.size()) {
    for (int64_t i = 0; i < (int64_t)deg.size(); ++i) {
      int64_t b = bucket(key[i]);
      int64_t h = head[b];
      next[i] = h;
      prev[i] = -1;
      if (h >= 0) prev[h] = i;
      head[b] = i;
    }
  }

  int64_t pop() {
    while (head[minlist] < 0) ++minlist;
    int64_t i = head[minlist];
    int64_t nx = next[i];
    head[minlist] = nx;
    if (nx >= 0) prev[nx] = -1;
    --nitems;
    return i;
  }

  void move(int64_t i, int64_t newkey, int64_t oldb, int64_t newb) {
    int64_t p = prev[i], nx = next[i];
    if (nx >= 0) prev[nx] = p;
    if (p >= 0)
      next[p] = nx;
    else
      head[oldb] = nx;
    int64_t h = head[newb];
    if (h >= 0) prev[h] = i;
    head[newb] = i;
    prev[i] = -1;
    next[i] = h;
    key[i] = newkey;
  }

  void dec(int64_t i) {
    int64_t ob = bucket(key[i]), nb = bucket(key[i] - 1);
    if (nb != ob) {
      move(i, key[i] - 1, ob, nb);
      if (nb < minlist) minlist = nb;
    } else {
      key[i] -= 1;
    }
  }

  void inc(int64_t i) {
    int64_t ob = bucket(key[i]), nb = bucket(key[i] + 1);
    if (nb != ob)
      move(i, key[i] + 1, ob, nb);
    else
      key[i] += 1;
  }
};

// Twin of every CSC entry (src/graphUtils.jl:49-55): flip[x] is the entry
// (c, r) of the entry x = (r, c).  The entries of column r sorted by row
// pair up, in order, with the entries of row r met while scanning columns
// c = 0..n-1, because both sequences run over c ascending.  Returns "" or
// the reason the matrix is not a valid adjacency (structurally asymmetric,
// nonzero diagonal).
std::string csc_twins(int64_t n, const int64_t* colptr, const int64_t* rowval, int64_t b,
                      std::vector<int64_t>& flip) {
  const int64_t m = colptr[n] - b;
  std::vector<int64_t> byrow(m);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t lo = colptr[i] - b, hi = colptr[i + 1] - b;
    for (int64_t k = lo; k < hi; ++k) {
      if (rowval[k] - b == i) return "nonzero diagonal entry";
      byrow[k] = k;
    }
    std::sort(byrow.begin() + lo, byrow.begin() + hi,
              [&](int64_t x, int64_t y) { return rowval[x] < rowval[y]; });
  }
  flip.assign(m, -1);
  std::vector<int64_t> cursor(n);
  for (int64_t i = 0; i < n; ++i) cursor[i] = colptr[i] - b;
  for (int64_t c = 0; c < n; ++c) {
    for (int64_t t = colptr[c] - b; t < colptr[c + 1] - b; ++t) {
      const int64_t x = byrow[t];  // entry (row r, col c), r ascending
      const int64_t r = rowval[x] - b;
      if (cursor[r] >= colptr[r + 1] - b) return "adjacency matrix is not structurally symmetric";
      const int64_t y = byrow[cursor[r]++];  // entry (row ?, col r)
      if (rowval[y] - b != c) return "adjacency matrix is not structurally symmetric";
      flip[x] = y;
    }
  }
  return "";
}

}  // namespace

struct lapfactor_ldli {
  int64_t n = 0;
  std::vector<int64_t> col, colptr, rowval;  // 0-based
  std::vector<double> fval, d;
};

extern "C" {

const char* lapfactor_last_error(void) { return g_err.c_str(); }

int64_t lapfactor_n(const lapfactor_ldli* f) { return f ? f->n : -1; }
int64_t lapfactor_nnz(const lapfactor_ldli* f) {
  return f ? (int64_t)f->fval.size() : -1;
}

void lapfactor_free(lapfactor_ldli* f) { delete f; }

int lapfactor_export(const lapfactor_ldli* f, int index_base, int64_t* col,
                     int64_t* colptr, int64_t* rowval, double* fval,
                     double* d) {
  if (!f) return fail(-1, "lapfactor_export: null factor");
  if (index_base != 0 && index_base != 1)
    return fail(-1, "lapfactor_export: index_base must be 0 or 1");
  const int64_t b = index_base;
  const int64_t n = f->n;
  const int64_t nnz = (int64_t)f->fval.size();
  if (col)
    for (int64_t i = 0; i < n - 1; ++i) col[i] = f->col[i] + b;
  if (colptr)
    for (int64_t i = 0; i < n; ++i) colptr[i] = f->colptr[i] + b;
  if (rowval)
    for (int64_t i = 0; i < nnz; ++i) rowval[i] = f->rowval[i] + b;
  if (fval) std::memcpy(fval, f->fval.data(), sizeof(double) * nnz);
  if (d) std::memcpy(d, f->d.data(), sizeof(double) * n);
  g_err.clear();
  return 0;
}

int lapfactor_build(int64_t n, const int64_t* colptr_in,
                    const int64_t* rowval_in, const double* nzval,
                    int index_base, uint64_t seed, lapfactor_ldli** out) {
  if (!out) return fail(-1, "lapfactor_build: null output handle");
  *out = nullptr;
  if (n < 1) return fail(-1, "lapfactor_build: n must be >= 1");
  if (index_base != 0 && index_base != 1)
    return fail(-1, "lapfactor_build: index_base must be 0 or 1");
  if (!colptr_in || (!rowval_in && n > 1) || (!nzval && n > 1))
    return fail(-1, "lapfactor_build: null input array");
  const int64_t b = index_base;
  const int64_t m = colptr_in[n] - b;
  if (colptr_in[0] - b != 0 || m < 0)
    return fail(-1, "lapfactor_build: malformed colptr");
  for (int64_t i = 0; i < n; ++i)
    if (colptr_in[i + 1] < colptr_in[i])
      return fail(-1, "lapfactor_build: colptr not monotone");
  for (int64_t k = 0; k < m; ++k) {
    if (nzval[k] < 0)
      return fail(-2, "Adjacency matrix can not have negative edge weights");
    int64_t r = rowval_in[k] - b;
    if (r < 0 || r >= n) return fail(-1, "lapfactor_build: row out of range");
  }

  lapfactor_ldli* F = new (std::nothrow) lapfactor_ldli();
  if (!F) return fail(-3, "lapfactor_build: out of memory");
  F->n = n;
  F->col.assign(n > 1 ? n - 1 : 0, 0);
  F->colptr.assign(n, 0);
  F->d.assign(n, 0.0);
  if (n == 1) {
    *out = F;
    g_err.clear();
    return 0;
  }

  try {
    // ---- LLmatp (src/approxChol.jl:102-136) --------------------------------
    std::vector<Edge> E(m);
    std::vector<int64_t> head(n), deg(n);
    for (int64_t i = 0; i < n; ++i) {
      int64_t lo = colptr_in[i] - b, hi = colptr_in[i + 1] - b;
      if (hi == lo) {
        delete F;
        return fail(-4, "lapfactor_build: isolated vertex (graph must be "
                        "connected; split components first)");
      }
      deg[i] = hi - lo;
      // the first stored entry is the list tail; later entries are pushed in
      // front, so the head is the last stored entry
      E[lo] = Edge{rowval_in[lo] - b, nzval[lo], lo, lo};
      for (int64_t k = lo + 1; k < hi; ++k)
        E[k] = Edge{rowval_in[k] - b, nzval[k], k - 1, k};
      head[i] = hi - 1;
    }
    {
      std::vector<int64_t> flip;
      const std::string why = csc_twins(n, colptr_in, rowval_in, b, flip);
      if (!why.empty()) {
        delete F;
        g_err = "lapfactor_build: " + why;
        return -1;
      }
      for (int64_t k = 0; k < m; ++k) E[k].rev = flip[k];
    }

    // ---- elimination (src/approxChol.jl:864-977) ---------------------------
    DegreeQueue pq(deg);
    Rng rng(seed);
    std::vector<int64_t> cs;   // column space: edge indices
    std::vector<double> cum, vals;
    cs.reserve(64);
    F->rowval.reserve((size_t)(1.1 * (double)m));
    F->fval.reserve((size_t)(1.1 * (double)m));

    int64_t it = 0;
    while (it < n - 1) {
      const int64_t i = pq.pop();
      F->col[it] = i;
      F->colptr[it] = (int64_t)F->fval.size();
      ++it;

      // get_ll_col (src/approxChol.jl:354-385): live entries of i's list
      cs.clear();
      int64_t e = head[i];
      while (true) {
        if (E[e].val > 0) cs.push_back(e);
        if (E[e].next == e) break;
        e = E[e].next;
      }
      if (cs.empty()) {
        delete F;
        return fail(-4, "lapfactor_build: vertex with no live edges "
                        "(disconnected graph)");
      }

      // compressCol! (src/approxChol.jl:415-449): merge multi-edges by row,
      // then order by weight (ties: ascending row)
      stable_sort_by(cs, cs.size(), [&](int64_t x) { return E[x].row; });
      size_t len = 0;
      int64_t currow = -1;
      for (size_t t = 0; t < cs.size(); ++t) {
        int64_t x = cs[t];
        if (E[x].row != currow) {
          currow = E[x].row;
          cs[len++] = x;
        } else {
          E[cs[len - 1]].val += E[x].val;
          E[E[x].rev].val = 0.0;
          pq.dec(currow);
        }
      }
      cs.resize(len);
      stable_sort_by(cs, len, [&](int64_t x) { return E[x].val; });

      vals.resize(len);
      cum.resize(len);
      double csum = 0.0;
      for (size_t t = 0; t < len; ++t) {
        vals[t] = E[cs[t]].val;
        csum += vals[t];
        cum[t] = csum;
      }
      double wdeg = csum;
      double colScale = 1.0;

      for (size_t t = 0; t + 1 < len; ++t) {
        const int64_t ll = cs[t];
        const double w = vals[t] * colScale;
        const int64_t j = E[ll].row;
        const int64_t revj = E[ll].rev;
        const double f = w / wdeg;
        vals[t] = 0.0;

        // sample k among later entries proportional to original weight:
        // first index with cum >= r (searchsortedfirst)
        const double r = rng.uniform() * (csum - cum[t]) + cum[t];
        size_t koff = (size_t)(std::lower_bound(cum.begin(), cum.end(), r) -
                               cum.begin());
        if (koff >= len) koff = len - 1;
        const int64_t k = E[cs[koff]].row;
        pq.inc(k);

        const double newEdgeVal = f * (1.0 - f) * wdeg;

        // row k in column j (reuse the twin entry)
        E[revj].row = k;
        E[revj].val = newEdgeVal;
        E[revj].rev = ll;
        // row j in column k (move ll to the front of k's list)
        const int64_t khead = head[k];
        head[k] = ll;
        E[ll].next = khead;
        E[ll].rev = revj;
        E[ll].val = newEdgeVal;
        E[ll].row = j;

        colScale *= (1.0 - f);
        wdeg *= (1.0 - f) * (1.0 - f);

        F->rowval.push_back(j);
        F->fval.push_back(f);
      }

      const int64_t ll = cs[len - 1];
      const double w = vals[len - 1] * colScale;
      const int64_t j = E[ll].row;
      if (it < n - 1) pq.dec(j);
      E[E[ll].rev].val = 0.0;
      F->rowval.push_back(j);
      F->fval.push_back(1.0);
      F->d[i] = w;
    }
    F->colptr[n - 1] = (int64_t)F->fval.size();
  } catch (const std::bad_alloc&) {
    delete F;
    return fail(-3, "lapfactor_build: out of memory");
  }

  *out = F;
  g_err.clear();
  return 0;
}

int lapfactor_build_split(int64_t n, const int64_t* colptr_in, const int64_t* rowval_in,
                          const double* nzval, int index_base, int64_t split, int64_t merge,
                          uint64_t seed, lapfactor_ldli** out) {
  if (!out) return fail(-1, "lapfactor_build_split: null output handle");
  *out = nullptr;
  if (n < 1) return fail(-1, "lapfactor_build_split: n must be >= 1");
  if (split < 1) return fail(-1, "lapfactor_build_split: split must be >= 1");
  // merge < 1: every copy is kept (avgCol!, src/approxChol.jl:497-533, the
  // split-only elimination of src/approxChol.jl:1162-1339)
  if (merge < 1) merge = INT64_MAX;
  if (index_base != 0 && index_base != 1)
    return fail(-1, "lapfactor_build_split: index_base must be 0 or 1");
  if (!colptr_in || (!rowval_in && n > 1) || (!nzval && n > 1))
    return fail(-1, "lapfactor_build_split: null input array");
  const int64_t b = index_base;
  const int64_t m = colptr_in[n] - b;
  if (colptr_in[0] - b != 0 || m < 0) return fail(-1, "lapfactor_build_split: malformed colptr");
  for (int64_t i = 0; i < n; ++i)
    if (colptr_in[i + 1] < colptr_in[i]) return fail(-1, "lapfactor_build_split: colptr not monotone");
  for (int64_t k = 0; k < m; ++k) {
    if (nzval[k] < 0) return fail(-2, "Adjacency matrix can not have negative edge weights");
    const int64_t r = rowval_in[k] - b;
    if (r < 0 || r >= n) return fail(-1, "lapfactor_build_split: row out of range");
  }
  for (int64_t i = 0; i < n; ++i)
    if (colptr_in[i + 1] == colptr_in[i])
      return fail(-4, "lapfactor_build_split: isolated vertex (graph must be connected; "
                      "split components first)");
  lapfactor_ldli* F = new (std::nothrow) lapfactor_ldli();
  if (!F) return fail(-3, "lapfactor_build_split: out of memory");
  F->n = n;
  F->col.assign(n > 1 ? n - 1 : 0, 0);
  F->colptr.assign(n, 0);
  F->d.assign(n, 0.0);
  if (n == 1) {
    *out = F;
    g_err.clear();
    return 0;
  }
  try {
    // ---- LLmatp(a, split) (src/approxChol.jl:138-186): every entry becomes
    // `split` copies of weight w / split, copy c of (r, c0) twinned with copy
    // c of (c0, r); slot of copy c of CSC entry k = k * split + c
    std::vector<int64_t> flip;
    const std::string why = csc_twins(n, colptr_in, rowval_in, b, flip);
    if (!why.empty()) {
      delete F;
      g_err = "lapfactor_build_split: " + why;
      return -1;
    }
    std::vector<Edge> E((size_t)(m * split));
    std::vector<int64_t> head(n), deg(n);
    for (int64_t i = 0; i < n; ++i) {
      const int64_t lo = colptr_in[i] - b, hi = colptr_in[i + 1] - b;
      deg[i] = split * (hi - lo);
      // list order as the reference links it: the first entry's last copy
      // is the tail, then every copy pushed in front in storage order
      int64_t prev = lo * split + split - 1;
      E[prev] = Edge{rowval_in[lo] - b, nzval[lo] / (double)split, prev, 0};
      for (int64_t k = lo; k < hi; ++k)
        for (int64_t c = 0; c < split; ++c) {
          const int64_t x = k * split + c;
          if (k == lo && c == split - 1) continue;
          E[x] = Edge{rowval_in[k] - b, nzval[k] / (double)split, prev, 0};
          prev = x;
        }
      head[i] = prev;
    }
    for (int64_t k = 0; k < m; ++k)
      for (int64_t c = 0; c < split; ++c) E[k * split + c].rev = flip[k] * split + c;

    // ---- approxChol(::LLmatp, split, merge) (src/approxChol.jl:979-1157;
    // with merge < 1, approxChol(::LLmatp, split), src/approxChol.jl:1162-1339)
    DegreeQueue pq(deg, split);
    Rng rng(seed);
    std::vector<int64_t> cs;
    std::vector<double> vals, crev;  // crev[t] = sum of vals[t..len)
    F->rowval.reserve((size_t)m);
    F->fval.reserve((size_t)m);
    int64_t it = 0;
    while (it < n - 1) {
      const int64_t i = pq.pop();
      F->col[it] = i;
      F->colptr[it] = (int64_t)F->fval.size();
      ++it;
      // get_ll_col (src/approxChol.jl:354-385): live entries
      cs.clear();
      for (int64_t e = head[i];; e = E[e].next) {
        if (E[e].val > 0) cs.push_back(e);
        if (E[e].next == e) break;
      }
      if (cs.empty()) {
        delete F;
        return fail(-4, "lapfactor_build_split: vertex with no live edges (disconnected graph)");
      }
      // compressAvgCol! (src/approxChol.jl:541-603): a row's copies get the
      // mean weight (both directions); rows with more than `merge` copies
      // keep `merge` copies carrying the whole weight, the rest are dropped;
      // then order by (weight, row)
      stable_sort_by(cs, cs.size(), [&](int64_t x) { return E[x].row; });
      size_t len = 0;
      for (size_t a0 = 0; a0 < cs.size();) {
        size_t a1 = a0;
        double sum = 0.0;
        while (a1 < cs.size() && E[cs[a1]].row == E[cs[a0]].row) sum += E[cs[a1++]].val;
        const int64_t num = (int64_t)(a1 - a0);
        const int64_t keep = num <= merge ? num : merge;
        const double v = sum / (double)keep;
        for (size_t t = a0; t < a1; ++t) {
          const int64_t x = cs[t];
          if ((int64_t)(t - a0) < keep) {
            E[x].val = v;
            E[E[x].rev].val = v;
            cs[len++] = x;
          } else {
            E[E[x].rev].val = 0.0;
            pq.dec(E[cs[a0]].row);
          }
        }
        a0 = a1;
      }
      cs.resize(len);
      stable_sort_by(cs, len, [&](int64_t x) { return std::make_pair(E[x].val, E[x].row); });
      vals.resize(len);
      crev.assign(len + 1, 0.0);
      double csum = 0.0;
      for (size_t t = 0; t < len; ++t) {
        vals[t] = E[cs[t]].val;
        csum += vals[t];
      }
      for (size_t t = len; t-- > 0;) crev[t] = crev[t + 1] + vals[t];
      double wdeg = csum, colScale = 1.0, final_d = 0.0;
      size_t next_edge = 0;  // first entry after the current row's copies
      for (size_t t = 0; t + 1 < len; ++t) {
        const bool first_copy = t == next_edge;
        while (next_edge < len && E[cs[t]].row == E[cs[next_edge]].row) ++next_edge;
        if (next_edge == len) {
          // the last row's copies: no samples, their weight goes to d
          for (; t + 1 < len; ++t) {
            final_d += vals[t] * colScale;
            if (it < n - 1) pq.dec(E[cs[t]].row);
            E[E[cs[t]].rev].val = 0.0;
          }
          break;
        }
        const int64_t ll = cs[t];
        const double w = vals[t] * colScale;
        const int64_t j = E[ll].row;
        const int64_t revj = E[ll].rev;
        const double f = w / wdeg;
        vals[t] = 0.0;
        // k among the entries after this row's copies, by weight: the last
        // index with crev >= r (searchsortedlast on the decreasing sums)
        const double r = rng.uniform() * crev[next_edge];
        size_t lo = next_edge, hi = len - 1;
        while (lo < hi) {
          const size_t mid = (lo + hi + 1) / 2;
          if (crev[mid] >= r)
            lo = mid;
          else
            hi = mid - 1;
        }
        const int64_t k = E[cs[lo]].row;
        pq.inc(k);
        const double newEdgeVal = E[ll].val * crev[next_edge] / csum;
        E[revj].row = k;
        E[revj].val = newEdgeVal;
        E[revj].rev = ll;
        const int64_t khead = head[k];
        head[k] = ll;
        E[ll].next = khead;
        E[ll].rev = revj;
        E[ll].val = newEdgeVal;
        E[ll].row = j;
        // one factor entry per row: the copies' f add up
        if (first_copy) {
          F->rowval.push_back(j);
          F->fval.push_back(f);
        } else {
          F->fval.back() += f;
        }
        if (t + 1 == next_edge) {
          const double ft = F->fval.back();
          colScale *= (1.0 - ft);
          wdeg *= (1.0 - ft) * (1.0 - ft);
        }
      }
      const int64_t ll = cs[len - 1];
      final_d += vals[len - 1] * colScale;
      const int64_t j = E[ll].row;
      if (it < n - 1) pq.dec(j);
      E[E[ll].rev].val = 0.0;
      F->rowval.push_back(j);
      F->fval.push_back(1.0);
      F->d[i] = final_d;
    }
    F->colptr[n - 1] = (int64_t)F->fval.size();
  } catch (const std::bad_alloc&) {
    delete F;
    return fail(-3, "lapfactor_build_split: out of memory");
  }
  *out = F;
  g_err.clear();
  return 0;
}

int lapfactor_build_ordered(int64_t n, const int64_t* colptr_in, const int64_t* rowval_in,
                            const double* nzval, int index_base, const int64_t* perm,
                            uint64_t seed, lapfactor_ldli** out) {
  if (!out) return fail(-1, "lapfactor_build_ordered: null output handle");
  *out = nullptr;
  if (n < 1) return fail(-1, "lapfactor_build_ordered: n must be >= 1");
  if (index_base != 0 && index_base != 1)
    return fail(-1, "lapfactor_build_ordered: index_base must be 0 or 1");
  if (!colptr_in || (!rowval_in && n > 1) || (!nzval && n > 1))
    return fail(-1, "lapfactor_build_ordered: null input array");
  const int64_t b = index_base;
  const int64_t m = colptr_in[n] - b;
  if (colptr_in[0] - b != 0 || m < 0) return fail(-1, "lapfactor_build_ordered: malformed colptr");
  for (int64_t i = 0; i < n; ++i)
    if (colptr_in[i + 1] < colptr_in[i])
      return fail(-1, "lapfactor_build_ordered: colptr not monotone");
  for (int64_t k = 0; k < m; ++k) {
    if (nzval[k] < 0) return fail(-2, "Adjacency matrix can not have negative edge weights");
    const int64_t r = rowval_in[k] - b;
    if (r < 0 || r >= n) return fail(-1, "lapfactor_build_ordered: row out of range");
  }
  // pos[v]: elimination position of vertex v (invperm of the order)
  std::vector<int64_t> order(n), pos(n, -1);
  for (int64_t t = 0; t < n; ++t) {
    const int64_t v = perm ? perm[t] - b : t;
    if (v < 0 || v >= n || pos[v] >= 0)
      return fail(-1, "lapfactor_build_ordered: perm is not a permutation");
    order[t] = v;
    pos[v] = t;
  }
  for (int64_t i = 0; i < n; ++i)
    if (colptr_in[i + 1] == colptr_in[i])
      return fail(-4, "lapfactor_build_ordered: isolated vertex (graph must be connected; "
                      "split components first)");

  lapfactor_ldli* F = new (std::nothrow) lapfactor_ldli();
  if (!F) return fail(-3, "lapfactor_build_ordered: out of memory");
  F->n = n;
  F->col.assign(n > 1 ? n - 1 : 0, 0);
  F->colptr.assign(n, 0);
  F->d.assign(n, 0.0);
  if (n == 1) {
    *out = F;
    g_err.clear();
    return 0;
  }
  try {
    // LLMatOrd (src/approxChol.jl:203-231, 265-296): every edge once, in
    // the list of its earlier-eliminated end, as (later position, weight);
    // an entry's slot is reused for the edge its elimination creates
    struct Ent {
      int64_t row;
      int64_t next;  // -1 ends a list
      double val;
    };
    std::vector<Ent> E;
    E.reserve((size_t)(m / 2 + 1));
    std::vector<int64_t> head(n, -1);
    for (int64_t i0 = 0; i0 < n; ++i0) {
      const int64_t i = pos[i0];
      for (int64_t k = colptr_in[i0] - b; k < colptr_in[i0 + 1] - b; ++k) {
        const int64_t j = pos[rowval_in[k] - b];
        if (j == i) {
          delete F;
          return fail(-1, "lapfactor_build_ordered: nonzero diagonal entry");
        }
        if (i < j) {
          E.push_back(Ent{j, head[i], nzval[k]});
          head[i] = (int64_t)E.size() - 1;
        }
      }
    }
    Rng rng(seed);
    struct Col {
      int64_t row, slot;
      double val;
    };
    std::vector<Col> cs;
    std::vector<double> cum;
    F->rowval.reserve((size_t)m);
    F->fval.reserve((size_t)m);
    // approxChol(::LLMatOrd) (src/approxChol.jl:758-862), in position space
    for (int64_t i = 0; i + 1 < n; ++i) {
      F->col[i] = order[i];
      F->colptr[i] = (int64_t)F->fval.size();
      cs.clear();
      for (int64_t e = head[i]; e >= 0; e = E[e].next) cs.push_back(Col{E[e].row, e, E[e].val});
      if (cs.empty()) {
        delete F;
        return fail(-4, "lapfactor_build_ordered: vertex with no later neighbour (disconnected graph)");
      }
      // compressCol! (src/approxChol.jl:451-495): merge by row (the merged
      // entry keeps the first slot), then order by weight (ties: row)
      std::stable_sort(cs.begin(), cs.end(), [](const Col& x, const Col& y) { return x.row < y.row; });
      size_t len = 0;
      for (size_t t = 0; t < cs.size(); ++t) {
        if (len && cs[len - 1].row == cs[t].row)
          cs[len - 1].val += cs[t].val;
        else
          cs[len++] = cs[t];
      }
      cs.resize(len);
      std::stable_sort(cs.begin(), cs.end(), [](const Col& x, const Col& y) { return x.val < y.val; });
      cum.resize(len);
      double csum = 0.0;
      for (size_t t = 0; t < len; ++t) {
        csum += cs[t].val;
        cum[t] = csum;
      }
      double wdeg = csum, colScale = 1.0;
      for (size_t t = 0; t + 1 < len; ++t) {
        const double w = cs[t].val * colScale;
        const int64_t j = cs[t].row;
        const double f = w / wdeg;
        const double r = rng.uniform() * (csum - cum[t]) + cum[t];
        size_t koff = (size_t)(std::lower_bound(cum.begin(), cum.end(), r) - cum.begin());
        if (koff >= len) koff = len - 1;
        const int64_t k = cs[koff].row;
        const double newEdgeVal = w * (1.0 - f);
        // edge (j, k) in the list of its earlier end, in this entry's slot
        const int64_t lo = j < k ? j : k, hi = j < k ? k : j;
        E[cs[t].slot] = Ent{hi, head[lo], newEdgeVal};
        head[lo] = cs[t].slot;
        colScale *= (1.0 - f);
        wdeg = wdeg - 2.0 * w + w * w / wdeg;
        F->rowval.push_back(order[j]);
        F->fval.push_back(f);
      }
      F->rowval.push_back(order[cs[len - 1].row]);
      F->fval.push_back(1.0);
      F->d[order[i]] = cs[len - 1].val * colScale;
    }
    F->colptr[n - 1] = (int64_t)F->fval.size();
  } catch (const std::bad_alloc&) {
    delete F;
    return fail(-3, "lapfactor_build_ordered: out of memory");
  }
  *out = F;
  g_err.clear();
  return 0;
}

}  // extern "C"
