// Test driver for the plan file (lapgpu_planio.cpp).
//   plan_io_check save <dir> <file>  builds the plan of the graph + factor in
//                                    <dir> (the arrays of plan_hash.cpp, with
//                                    a dense block and the level groups),
//                                    writes it, loads it back and prints
//                                    "equal 1" when every array and scalar
//                                    of the loaded plan equals the built one
//   plan_io_check load <file>        prints "ok n <n>" or "error <message>"
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
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

template <class T>
static bool same(const std::vector<T>& a, const std::vector<T>& b) {
  return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), sizeof(T) * a.size()) == 0);
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const std::string mode = argv[1];
  std::string err;
  if (mode == "load") {
    Plan P;
    if (!load_plan(argv[2], P, err)) {
      printf("error %s\n", err.c_str());
      return 0;
    }
    printf("ok n %lld\n", (long long)P.n);
    return 0;
  }
  if (mode != "save" || argc < 4) return 2;
  const std::string d = argv[2];
  auto cp = read_all<int64_t>(d + "/cp"), rv = read_all<int64_t>(d + "/rv");
  auto nz = read_all<double>(d + "/nz");
  auto col = read_all<int64_t>(d + "/col"), fcp = read_all<int64_t>(d + "/fcp"),
       frv = read_all<int64_t>(d + "/frv");
  auto fv = read_all<double>(d + "/fv"), dd = read_all<double>(d + "/d");
  const int64_t n = (int64_t)cp.size() - 1;
  HostCsr L;
  if (!csc_to_csr(n, cp.data(), rv.data(), nz.data(), 0, L, err)) {
    printf("error %s\n", err.c_str());
    return 1;
  }
  Plan P;
  DenseModel dm;
  dm.max_k = kTailCap;
  if (!build_plan(L, (int64_t)fv.size(), col.data(), fcp.data(), frv.data(), fv.data(), dd.data(),
                  0, 24, P, err, dm)) {
    printf("error %s\n", err.c_str());
    return 1;
  }
  build_blocks(P, 16, 4.0, 1 << 22);
  if (!save_plan(P, argv[3], err)) {
    printf("error %s\n", err.c_str());
    return 1;
  }
  Plan Q;
  if (!load_plan(argv[3], Q, err)) {
    printf("error %s\n", err.c_str());
    return 1;
  }
  const TailProgram &A = P.tail, &B = Q.tail;
  const DenseBlock &DA = P.dense, &DB = Q.dense;
  const bool eq =
      P.n == Q.n && P.nnzF == Q.nnzF && P.nlev == Q.nlev && P.tail_level == Q.tail_level &&
      same(P.perm, Q.perm) && same(P.iperm, Q.iperm) && same(P.Lptr, Q.Lptr) &&
      same(P.Lidx, Q.Lidx) && same(P.Lval, Q.Lval) && same(P.Fptr, Q.Fptr) &&
      same(P.Fidx, Q.Fidx) && same(P.Fw, Q.Fw) && same(P.Bptr, Q.Bptr) && same(P.Bidx, Q.Bidx) &&
      same(P.Bw, Q.Bw) && same(P.s, Q.s) && same(P.lvl, Q.lvl) && A.T0 == B.T0 && A.NT == B.NT &&
      A.NT1 == B.NT1 && A.elog == B.elog && A.npart == B.npart && same(A.fsteps, B.fsteps) &&
      same(A.bsteps, B.bsteps) && same(A.fslots, B.fslots) && same(A.bslots, B.bslots) &&
      same(A.Eptr, B.Eptr) && same(A.Eidx, B.Eidx) && same(A.Ew, B.Ew) && DA.T1 == DB.T1 &&
      DA.K == DB.K && DA.level == DB.level && DA.elog == DB.elog && DA.npart == DB.npart &&
      same(DA.fsteps, DB.fsteps) && same(DA.bsteps, DB.bsteps) && same(DA.fslots, DB.fslots) &&
      same(DA.bslots, DB.bslots) && same(P.blocks.lev, Q.blocks.lev) &&
      same(P.blocks.fptr, Q.blocks.fptr) && same(P.blocks.fidx, Q.blocks.fidx) &&
      same(P.blocks.fw, Q.blocks.fw) && same(P.blocks.bptr, Q.blocks.bptr) &&
      same(P.blocks.bidx, Q.blocks.bidx) && same(P.blocks.bw, Q.blocks.bw);
  printf("equal %d n %lld K %lld groups %zu\n", eq ? 1 : 0, (long long)P.n, (long long)DA.K,
         P.blocks.lev.empty() ? (size_t)0 : P.blocks.lev.size() - 1);
  return 0;
}
