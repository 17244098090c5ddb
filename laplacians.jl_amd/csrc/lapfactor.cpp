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

struct Edge {
  int64_t row;
  double val;
  int64_t next;  // next edge in the owning vertex's list; == self at the tail
  int64_t rev;   // twin entry in the row vertex's list
};

// Bucketed approximate priority queue on integer keys.
struct DegreeQueue {
  int64_t n;
  std::vector<int64_t> prev, next, key;  // per vertex; -1 = none
  std::vector<int64_t> head;             // per bucket
  int64_t minlist = 1;
  int64_t nitems;

  int64_t bucket(int64_t k) const {
    if (k <= n) return k;
    int64_t b = n + k / n;
    return b < 2 * n + 1 ? b : 2 * n + 1;
  }

  explicit DegreeQueue(const std::vector<int64_t>& deg)
      : n((int64_t)deg.size()),
        prev(deg.size(), -1),
        next(deg.size(), -1),
        key(deg),
        head(2 * deg.size() + 2, -1),
        nitems((int64_t)deg.size()) {
    for (int64_t i = 0; i < n; ++i) {
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
    // twin index (src/graphUtils.jl:49-55): the entries of column r sorted
    // by row pair up, in order, with the entries (row r) met while scanning
    // columns c = 0..n-1, because both sequences run over c ascending.
    {
      std::vector<int64_t> byrow(m);
      for (int64_t i = 0; i < n; ++i) {
        int64_t lo = colptr_in[i] - b, hi = colptr_in[i + 1] - b;
        for (int64_t k = lo; k < hi; ++k) byrow[k] = k;
        std::sort(byrow.begin() + lo, byrow.begin() + hi,
                  [&](int64_t x, int64_t y) { return E[x].row < E[y].row; });
      }
      std::vector<int64_t> cursor(n);
      for (int64_t i = 0; i < n; ++i) cursor[i] = colptr_in[i] - b;
      for (int64_t c = 0; c < n; ++c) {
        for (int64_t t = colptr_in[c] - b; t < colptr_in[c + 1] - b; ++t) {
          const int64_t x = byrow[t];  // entry (row r, col c), r ascending
          const int64_t r = E[x].row;
          if (cursor[r] >= colptr_in[r + 1] - b) {
            delete F;
            return fail(-1, "lapfactor_build: adjacency matrix is not "
                            "structurally symmetric");
          }
          const int64_t y = byrow[cursor[r]++];  // entry (row ?, col r)
          if (E[y].row != c) {
            delete F;
            return fail(-1, "lapfactor_build: adjacency matrix is not "
                            "structurally symmetric");
          }
          E[x].rev = y;
        }
      }
    }
    for (int64_t i = 0; i < n; ++i)
      for (int64_t k = colptr_in[i] - b; k < colptr_in[i + 1] - b; ++k)
        if (E[k].row == i) {
          delete F;
          return fail(-1, "lapfactor_build: nonzero diagonal entry");
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
      std::stable_sort(cs.begin(), cs.end(), [&](int64_t x, int64_t y) {
        return E[x].row < E[y].row;
      });
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
      std::stable_sort(cs.begin(), cs.end(), [&](int64_t x, int64_t y) {
        return E[x].val < E[y].val;
      });

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

}  // extern "C"
