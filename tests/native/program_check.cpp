// Test driver: runs the slot programs the planner builds (the single-
// workgroup tail of a context, its dense block's own programs, and the
// whole-graph programs of the batched engine) on the host, lane by lane as
// the kernels do, and compares them with the sweeps written out directly
// over the level-order layout:
//   forward:  y0[v] = r[v] + sum_k Fw[k] y0[Fidx[k]]
//   backward: yF[v] = s[v] y0[v] + sum_k Bw[k] yF[Bidx[k]]
// Partial slots start as NaN, so a fix-up that reads a slot no chunk wrote
// shows up as a NaN error.
// Usage: program_check <dir> <ctx|wg> <dense_max> <det 0|1>
// Input: the raw arrays of plan_hash.cpp (la in CSC, the LDLinv).
// Prints: err_fwd err_bwd err_dense npart npart_dense lead2 lead3 steps T0 NT K
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include "lapgpu_plan.hpp"

using namespace lapgpu;

template <class T>
static std::vector<T> read_all(const std::string& f) {
  std::ifstream in(f, std::ios::binary | std::ios::ate);
  const size_t sz = (size_t)in.tellg();
  in.seekg(0);
  std::vector<T> v(sz / sizeof(T));
  in.read((char*)v.data(), (std::streamsize)sz);
  return v;
}

struct Counts {
  int64_t lead2 = 0, lead3 = 0, steps = 0;
};

// One program on y (local indices), as the kernels run it: per step, every
// lane's E products, the sum over each aligned segment of 2^seglog lanes
// into its last lane, and the lead lanes' adds or partial stores.
static void run(const std::vector<TailStep>& steps, const std::vector<TailSlot>& slots,
                std::vector<double>& y, Counts& cnt) {
  for (const TailStep& st : steps) {
    const int64_t count = st.meta & 0xffff;
    const int64_t E = int64_t(1) << ((st.meta >> 17) & 3);
    std::vector<double> v(count, 0.0);
    for (int64_t t = 0; t < count; ++t)
      for (int64_t e = 0; e < E; ++e) {
        const TailSlot& S = slots[st.off + e * count + t];
        if (S.w != 0.0) v[t] += S.w * y[S.src];
      }
    struct W {
      uint32_t row, lead;
      double v;
    };
    std::vector<W> writes;
    for (int64_t t = 0; t < count; ++t) {
      const uint32_t meta = slots[st.off + t].meta;
      const uint32_t lead = meta >> 29;
      if (lead == 0) continue;
      const int64_t seg = int64_t(1) << ((meta >> 26) & 7);
      double sum = 0.0;
      for (int64_t u = t - seg + 1; u <= t; ++u) sum += v[u];
      writes.push_back({meta & kRowMask, lead, sum});
      cnt.lead2 += lead == 2;
      cnt.lead3 += lead == 3;
    }
    for (const W& w : writes) {
      if (w.lead == 3)
        y[w.row] = w.v;
      else
        y[w.row] += w.v;
    }
    ++cnt.steps;
  }
}

static double rel(const std::vector<double>& a, const std::vector<double>& b, int64_t a0, int64_t b0,
                  int64_t len) {
  double num = 0.0, den = 0.0;
  for (int64_t i = 0; i < len; ++i) {
    const double d = a[a0 + i] - b[b0 + i];
    num = std::fmax(num, std::fabs(d));
    if (std::isnan(d)) return NAN;
    den = std::fmax(den, std::fabs(b[b0 + i]));
  }
  return den > 0 ? num / den : num;
}

int main(int argc, char** argv) {
  if (argc < 5) return 2;
  const std::string d = argv[1], mode = argv[2];
  const int64_t dense_max = std::stoll(argv[3]);
  const bool det = std::string(argv[4]) == "1";
  auto cp = read_all<int64_t>(d + "/cp"), rv = read_all<int64_t>(d + "/rv");
  auto nz = read_all<double>(d + "/nz");
  auto col = read_all<int64_t>(d + "/col"), fcp = read_all<int64_t>(d + "/fcp"),
       frv = read_all<int64_t>(d + "/frv");
  auto fv = read_all<double>(d + "/fv"), dd = read_all<double>(d + "/d");
  const int64_t n = (int64_t)cp.size() - 1;
  HostCsr L;
  std::string err;
  if (!csc_to_csr(n, cp.data(), rv.data(), nz.data(), 0, L, err)) {
    printf("error %s\n", err.c_str());
    return 1;
  }
  Plan P;
  DenseModel dm;
  dm.max_k = mode == "ctx" ? dense_max : 0;
  if (!build_plan(L, (int64_t)fv.size(), col.data(), fcp.data(), frv.data(), fv.data(), dd.data(),
                  0, 24, P, err, dm)) {
    printf("error %s\n", err.c_str());
    return 1;
  }
  bool ok = mode == "wg" ? build_tail(P, INT64_MAX, INT64_MAX, 0, 1024, DenseModel(), det)
                         : build_tail(P, 24, kTailCap, kTailMaxElog, kTailLanes, dm, det);
  if (!ok) {
    printf("error build_tail\n");
    return 1;
  }
  // the direct sweeps on a random r
  std::mt19937_64 g(7);
  std::normal_distribution<double> nd;
  std::vector<double> r(n), y0(n), yF(n);
  for (auto& x : r) x = nd(g);
  for (int64_t v = 0; v < n; ++v) {
    double a = r[v];
    for (int32_t k = P.Fptr[v]; k < P.Fptr[v + 1]; ++k) a += P.Fw[k] * y0[P.Fidx[k]];
    y0[v] = a;
  }
  for (int64_t v = n - 1; v >= 0; --v) {
    double a = P.s[v] * y0[v];
    for (int32_t k = P.Bptr[v]; k < P.Bptr[v + 1]; ++k) a += P.Bw[k] * yF[P.Bidx[k]];
    yF[v] = a;
  }
  const TailProgram& T = P.tail;
  const int64_t T0 = T.T0, NT = T.NT, NT1 = T.NT1, T1 = T0 + NT1;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  Counts cnt;
  // forward: y = r + the in-edges from outside the tail
  std::vector<double> y(NT + T.npart, nan);
  for (int64_t i = 0; i < NT; ++i) {
    double a = r[T0 + i];
    for (int32_t k = T.Eptr[i]; k < T.Eptr[i + 1]; ++k) a += T.Ew[k] * y0[T.Eidx[k]];
    y[i] = a;
  }
  run(T.fsteps, T.fslots, y, cnt);
  // [T0, T1) = y0; with a dense block [T1, n) holds its inputs
  // u[v] = r[v] + sum over the in-edges from below T1
  std::vector<double> u(n - T1);
  for (int64_t v = T1; v < n; ++v) {
    double a = r[v];
    for (int32_t k = P.Fptr[v]; k < P.Fptr[v + 1]; ++k)
      if (P.Fidx[k] < T1) a += P.Fw[k] * y0[P.Fidx[k]];
    u[v - T1] = a;
  }
  double ef = rel(y, y0, 0, T0, NT1);
  if (NT1 < NT) ef = std::fmax(ef, rel(y, u, NT1, 0, NT - NT1));
  // backward: [T0, T1) scaled, [T1, n) the block's results
  for (int64_t i = 0; i < NT1; ++i) y[i] = P.s[T0 + i] * y0[T0 + i];
  for (int64_t i = NT1; i < NT; ++i) y[i] = yF[T0 + i];
  for (int64_t i = NT; i < NT + T.npart; ++i) y[i] = nan;
  run(T.bsteps, T.bslots, y, cnt);
  const double eb = rel(y, yF, 0, T0, NT1);
  // the dense block's own programs (local to T1) on u
  double ed = 0.0;
  const DenseBlock& D = P.dense;
  if (D.K > 0) {
    std::vector<double> z(D.K + D.npart, nan);
    for (int64_t i = 0; i < D.K; ++i) z[i] = u[i];
    run(D.fsteps, D.fslots, z, cnt);
    ed = rel(z, y0, 0, T1, D.K);
    for (int64_t i = 0; i < D.K; ++i) z[i] *= P.s[T1 + i];
    run(D.bsteps, D.bslots, z, cnt);
    ed = std::fmax(ed, rel(z, yF, 0, T1, D.K));
  }
  printf("err_fwd %.3e err_bwd %.3e err_dense %.3e npart %lld npart_dense %lld lead2 %lld lead3 %lld steps %lld "
         "T0 %lld NT %lld K %lld\n",
         ef, eb, ed, (long long)T.npart, (long long)D.npart, (long long)cnt.lead2,
         (long long)cnt.lead3, (long long)cnt.steps, (long long)T0, (long long)NT, (long long)D.K);
  return 0;
}
