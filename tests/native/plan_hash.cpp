// Test driver: builds the device plan of one graph + factor on the host and
// prints a hash of every plan array (FNV-1a over the bytes), so that plans
// built with different thread counts can be compared bit for bit.
// With a second argument "root_s" it prints the diagonal scale of the
// never-eliminated vertex instead.
// Input: a directory with raw little-endian arrays cp, rv (int64), nz
// (float64) of la in CSC, and col, fcp, frv (int64), fv, d (float64) of the
// LDLinv, all 0-based.
#include <cstdint>
#include <cstdio>
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

template <class V>
static unsigned long long fnv(const V& v) {
  uint64_t x = 1469598103934665603ull;
  const unsigned char* c = (const unsigned char*)v.data();
  for (size_t i = 0; i < v.size() * sizeof(v[0]); ++i) {
    x ^= c[i];
    x *= 1099511628211ull;
  }
  return (unsigned long long)x;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string d = argv[1];
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
  if (!build_plan(L, (int64_t)fv.size(), col.data(), fcp.data(), frv.data(), fv.data(),
                  dd.data(), 0, 24, P, err)) {
    printf("error %s\n", err.c_str());
    return 1;
  }
  if (argc > 2 && std::string(argv[2]) == "root_s") {
    // the diagonal scale of the never-eliminated vertex, at its position
    std::vector<char> elim(n, 0);
    for (int64_t c : col) elim[c] = 1;
    for (int64_t v = 0; v < n; ++v)
      if (!elim[v]) printf("root %lld s %.17g\n", (long long)v, P.s[P.iperm[v]]);
    return 0;
  }
  build_blocks(P, 16, 4.0, 1 << 22);
#define H(v) printf("%s %zu %016llx\n", #v, P.v.size(), fnv(P.v))
  H(perm); H(Lptr); H(Lidx); H(Lval); H(Fptr); H(Fidx); H(Fw); H(Bptr); H(Bidx); H(Bw);
  H(s); H(lvl); H(blocks.lev); H(blocks.fptr); H(blocks.fidx); H(blocks.fw);
  H(blocks.bptr); H(blocks.bidx); H(blocks.bw); H(tail.fsteps); H(tail.fslots);
  H(tail.bsteps); H(tail.bslots); H(tail.Eptr); H(tail.Eidx); H(tail.Ew);
  printf("levels %lld groups %zu tail_vertices %lld elog %d\n", (long long)P.nlev,
         P.blocks.lev.size() - 1, (long long)P.tail.NT, P.tail.elog);
  return 0;
}
