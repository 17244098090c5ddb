// lapgpu_planio.cpp — the plan file: a context's whole host plan (level
// order, permuted Laplacian, both sweep layouts, tail and dense-block slot
// programs, level-group composite rows) written once and loaded instead of
// planning again.
//
// Layout (little-endian): the 8 bytes "LAPGPUPL", u32 version, u32 record
// count, then per record: u16 name length, the name, u8 element type
// (1 int32, 2 int64, 3 float64, 4 raw bytes), u32 element size, u64 element
// count, the elements; finally a u64 checksum: the fold of the header's
// and every record's header and payload hashes, in file order (payloads are
// hashed in 64 MiB pieces on several threads).  Records come in a fixed
// order.  Files are written to
// `path.tmp` and renamed, so a reader never sees a partial file; a file
// whose checksum, version, records, shapes or indices do not fit is refused.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <future>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "lapgpu_plan.hpp"

namespace lapgpu {

namespace {

constexpr char kMagic[8] = {'L', 'A', 'P', 'G', 'P', 'U', 'P', 'L'};
constexpr uint32_t kVersion = 1;

// Checksum of the byte stream: each piece written (a record header, a
// payload) is folded in 32-byte blocks by four multiply-xorshift lanes, its
// last bytes zero-padded with their count mixed in; the reader folds the
// same pieces.
struct Hasher {
  uint64_t h[4] = {0x9e3779b97f4a7c15ull, 0xbf58476d1ce4e5b9ull, 0x94d049bb133111ebull,
                   0x2545f4914f6cdd1dull};
  static uint64_t mix(uint64_t a, uint64_t w) {
    a = (a ^ w) * 0xff51afd7ed558ccdull;
    return a ^ (a >> 29);
  }
  void add(const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    size_t i = 0;
    for (; i + 32 <= n; i += 32) {
      uint64_t w[4];
      std::memcpy(w, c + i, 32);
      for (int l = 0; l < 4; ++l) h[l] = mix(h[l], w[l]);
    }
    uint64_t w[4] = {0, 0, 0, 0};
    if (n > i) std::memcpy(w, c + i, n - i);
    for (int l = 0; l < 4; ++l) h[l] = mix(h[l], w[l] ^ ((uint64_t)(n - i) << 56));
  }
  uint64_t value() const { return mix(mix(h[0], h[1]), mix(h[2], h[3])); }
};

// f(lo, hi) over [0, n) in pieces of `piece`, on up to the machine's
// threads; returns the pieces' results in order.
template <class R, class F>
std::vector<R> in_pieces(size_t n, size_t piece, F f) {
  const size_t np = n ? (n + piece - 1) / piece : 0;
  std::vector<R> out(np);
  const size_t nt = std::max<size_t>(1, std::min<size_t>(np, std::thread::hardware_concurrency()));
  std::vector<std::future<void>> fs;
  for (size_t t = 0; t < nt; ++t)
    fs.push_back(std::async(std::launch::async, [&, t] {
      for (size_t i = t; i < np; i += nt) out[i] = f(i * piece, std::min(n, (i + 1) * piece));
    }));
  for (auto& x : fs) x.get();
  return out;
}

// Hash of a payload: its 64 MiB pieces hashed in parallel, folded in order.
uint64_t payload_hash(const void* p, size_t n) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  const auto parts = in_pieces<uint64_t>(n, size_t(64) << 20, [&](size_t lo, size_t hi) {
    Hasher h;
    h.add(c + lo, hi - lo);
    return h.value();
  });
  Hasher h;
  h.add(&n, 8);
  if (!parts.empty()) h.add(parts.data(), 8 * parts.size());
  return h.value();
}

constexpr uint8_t kI32 = 1, kI64 = 2, kF64 = 3, kRaw = 4;
template <class T>
constexpr uint8_t type_of() {
  return std::is_same<T, int32_t>::value ? kI32 : std::is_same<T, double>::value ? kF64 : kRaw;
}

// The records of a plan in file order: f(name, vector) for arrays and
// f(name, int64) for scalars (the elog fields through te / de).
template <class Plan_, class I, class F>
void for_records(Plan_& P, I& te, I& de, F&& f) {
  f("n", P.n);
  f("nnzF", P.nnzF);
  f("perm", P.perm);
  f("iperm", P.iperm);
  f("Lptr", P.Lptr);
  f("Lidx", P.Lidx);
  f("Lval", P.Lval);
  f("Fptr", P.Fptr);
  f("Fidx", P.Fidx);
  f("Fw", P.Fw);
  f("Bptr", P.Bptr);
  f("Bidx", P.Bidx);
  f("Bw", P.Bw);
  f("s", P.s);
  f("lvl", P.lvl);
  f("nlev", P.nlev);
  f("tail_level", P.tail_level);
  auto& T = P.tail;
  f("tail.T0", T.T0);
  f("tail.NT", T.NT);
  f("tail.NT1", T.NT1);
  f("tail.elog", te);
  f("tail.npart", T.npart);
  f("tail.fsteps", T.fsteps);
  f("tail.bsteps", T.bsteps);
  f("tail.fslots", T.fslots);
  f("tail.bslots", T.bslots);
  f("tail.Eptr", T.Eptr);
  f("tail.Eidx", T.Eidx);
  f("tail.Ew", T.Ew);
  auto& D = P.dense;
  f("dense.T1", D.T1);
  f("dense.K", D.K);
  f("dense.level", D.level);
  f("dense.elog", de);
  f("dense.npart", D.npart);
  f("dense.fsteps", D.fsteps);
  f("dense.bsteps", D.bsteps);
  f("dense.fslots", D.fslots);
  f("dense.bslots", D.bslots);
  auto& B = P.blocks;
  f("blocks.lev", B.lev);
  f("blocks.fptr", B.fptr);
  f("blocks.fidx", B.fidx);
  f("blocks.fw", B.fw);
  f("blocks.bptr", B.bptr);
  f("blocks.bidx", B.bidx);
  f("blocks.bw", B.bw);
}

struct Counter {
  uint32_t n = 0;
  template <class T>
  void operator()(const char*, const T&) {
    ++n;
  }
};

struct Writer {
  std::FILE* f = nullptr;
  Hasher hs;
  bool ok = true;
  void put(const void* p, size_t n) {
    const uint64_t ph = payload_hash(p, n);
    hs.add(&ph, 8);
    if (n && std::fwrite(p, 1, n, f) != n) ok = false;
  }
  void rec(const char* name, uint8_t type, uint32_t esize, uint64_t n, const void* p) {
    unsigned char head[2 + 64 + 13];
    const uint16_t nl = (uint16_t)std::strlen(name);
    std::memcpy(head, &nl, 2);
    std::memcpy(head + 2, name, nl);
    head[2 + nl] = type;
    std::memcpy(head + 3 + nl, &esize, 4);
    std::memcpy(head + 7 + nl, &n, 8);
    put(head, 15 + (size_t)nl);
    put(p, (size_t)esize * (size_t)n);
  }
  void operator()(const char* name, const int64_t& x) { rec(name, kI64, 8, 1, &x); }
  template <class T>
  void operator()(const char* name, const std::vector<T>& v) {
    rec(name, type_of<T>(), sizeof(T), v.size(), v.data());
  }
};

struct Reader {
  std::FILE* f = nullptr;
  Hasher hs;
  std::string bad;  // the first record that did not fit
  bool raw(void* p, size_t n) { return n == 0 || std::fread(p, 1, n, f) == n; }
  bool get(void* p, size_t n) {
    if (!raw(p, n)) return false;
    const uint64_t ph = payload_hash(p, n);
    hs.add(&ph, 8);
    return true;
  }
  // one record header (hashed as the writer's one piece); true when it
  // names `name` with this type and element size
  bool head(const char* name, uint8_t type, uint32_t esize, uint64_t* n) {
    if (!bad.empty()) return false;
    unsigned char h[2 + 64 + 13];
    const size_t nl = std::strlen(name);
    uint16_t fl = 0;
    if (!raw(h, 2) || (std::memcpy(&fl, h, 2), fl != nl) || !raw(h + 2, nl + 13) ||
        std::memcmp(h + 2, name, nl) != 0) {
      bad = name;
      return false;
    }
    const uint64_t ph = payload_hash(h, 15 + nl);
    hs.add(&ph, 8);
    uint32_t es;
    std::memcpy(&es, h + 3 + nl, 4);
    std::memcpy(n, h + 7 + nl, 8);
    if (h[2 + nl] != type || es != esize) {
      bad = name;
      return false;
    }
    return true;
  }
  void operator()(const char* name, int64_t& x) {
    uint64_t n = 0;
    if (!head(name, kI64, 8, &n)) return;
    if (n != 1 || !get(&x, 8)) bad = name;
  }
  template <class T>
  void operator()(const char* name, std::vector<T>& v) {
    uint64_t n = 0;
    if (!head(name, type_of<T>(), sizeof(T), &n)) return;
    if (n > (uint64_t(1) << 40) / sizeof(T)) {
      bad = name;
      return;
    }
    v.resize((size_t)n);
    if (!get(v.data(), sizeof(T) * (size_t)n)) bad = name;
  }
};

// Every index of a slot program within [0, span) and every step's slots
// within the slot array.
bool program_in_bounds(const std::vector<TailStep>& steps, const std::vector<TailSlot>& slots,
                       int64_t span) {
  for (const TailStep& st : steps) {
    const int64_t count = st.meta & 0xffff;
    const int64_t E = int64_t(1) << ((st.meta >> 17) & 3);
    if (st.off < 0 || count % 64 != 0 || (int64_t)st.off + E * count > (int64_t)slots.size())
      return false;
  }
  for (const TailSlot& sl : slots)
    if ((int64_t)sl.src >= span || (int64_t)(sl.meta & kRowMask) >= span) return false;
  return true;
}

template <class T>
bool all_within(const std::vector<T>& v, int64_t lo, int64_t hi) {
  const auto ok = in_pieces<char>(v.size(), size_t(1) << 24, [&](size_t a, size_t b) -> char {
    for (size_t i = a; i < b; ++i)
      if ((int64_t)v[i] < lo || (int64_t)v[i] >= hi) return 0;
    return 1;
  });
  return std::all_of(ok.begin(), ok.end(), [](char c) { return c != 0; });
}

struct FileCloser {
  std::FILE* f;
  ~FileCloser() {
    if (f) std::fclose(f);
  }
};

}  // namespace

bool save_plan(const Plan& P, const std::string& path, std::string& err) {
  const int64_t te = P.tail.elog, de = P.dense.elog;
  Counter cnt;
  for_records(P, te, de, cnt);
  const std::string tmp = path + ".tmp";
  Writer w;
  w.f = std::fopen(tmp.c_str(), "wb");
  if (!w.f) {
    err = "save_plan: cannot write " + tmp;
    return false;
  }
  {
    FileCloser fc{w.f};
    std::vector<char> vb(size_t(1) << 22);
    std::setvbuf(w.f, vb.data(), _IOFBF, vb.size());
    w.put(kMagic, 8);
    const uint32_t ver = kVersion;
    w.put(&ver, 4);
    w.put(&cnt.n, 4);
    for_records(P, te, de, w);
    const uint64_t h = w.hs.value();
    if (std::fwrite(&h, 8, 1, w.f) != 1) w.ok = false;
    if (std::fflush(w.f) != 0) w.ok = false;
  }
  if (!w.ok) {
    std::remove(tmp.c_str());
    err = "save_plan: write failed: " + tmp;
    return false;
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) {
    std::remove(tmp.c_str());
    err = "save_plan: cannot rename " + tmp + " to " + path;
    return false;
  }
  return true;
}

bool load_plan(const std::string& path, Plan& P, std::string& err) {
  Reader r;
  r.f = std::fopen(path.c_str(), "rb");
  if (!r.f) {
    err = "load_plan: cannot read " + path;
    return false;
  }
  FileCloser fc{r.f};
  std::vector<char> vb(size_t(1) << 22);
  std::setvbuf(r.f, vb.data(), _IOFBF, vb.size());
  char magic[8];
  uint32_t ver = 0, cnt = 0;
  if (!r.get(magic, 8) || std::memcmp(magic, kMagic, 8) != 0) {
    err = "load_plan: " + path + " is not a plan file";
    return false;
  }
  if (!r.get(&ver, 4) || ver != kVersion) {
    err = "load_plan: plan file version " + std::to_string(ver) + ", expected " +
          std::to_string(kVersion);
    return false;
  }
  Plan Q;
  int64_t te = 0, de = 0;
  Counter want;
  for_records(Q, te, de, want);
  if (!r.get(&cnt, 4) || cnt != want.n) {
    err = "load_plan: " + path + " holds " + std::to_string(cnt) + " records, expected " +
          std::to_string(want.n);
    return false;
  }
  for_records(Q, te, de, r);
  if (!r.bad.empty()) {
    err = "load_plan: record " + r.bad + " missing, truncated or of the wrong type in " + path;
    return false;
  }
  uint64_t h = 0;
  char extra;
  if (std::fread(&h, 8, 1, r.f) != 1 || h != r.hs.value() || std::fread(&extra, 1, 1, r.f) != 0) {
    err = "load_plan: checksum mismatch (truncated or corrupted file): " + path;
    return false;
  }
  TailProgram& T = Q.tail;
  DenseBlock& D = Q.dense;
  LevelBlocks& B = Q.blocks;
  T.elog = (int)te;
  D.elog = (int)de;
  // the shapes and indices the device code relies on
  const int64_t n = Q.n;
  const bool shapes =
      n > 0 && (int64_t)Q.perm.size() == n && (int64_t)Q.iperm.size() == n &&
      (int64_t)Q.Lptr.size() == n + 1 && (int64_t)Q.Fptr.size() == n + 1 &&
      (int64_t)Q.Bptr.size() == n + 1 && (int64_t)Q.s.size() == n && Q.nlev >= 1 &&
      (int64_t)Q.lvl.size() == Q.nlev + 1 && Q.Lidx.size() == Q.Lval.size() &&
      (int64_t)Q.Lptr.back() == (int64_t)Q.Lidx.size() && Q.Fidx.size() == Q.Fw.size() &&
      (int64_t)Q.Fptr.back() == (int64_t)Q.Fidx.size() && Q.Bidx.size() == Q.Bw.size() &&
      (int64_t)Q.Bptr.back() == (int64_t)Q.Bidx.size() && T.T0 >= 0 && T.T0 + T.NT == n &&
      T.NT1 >= 0 && T.NT1 <= T.NT && T.NT <= kTailCap && (int64_t)T.Eptr.size() == T.NT + 1 &&
      T.Eidx.size() == T.Ew.size() && (int64_t)T.Eptr.back() == (int64_t)T.Eidx.size() &&
      T.elog >= 0 && T.elog <= kTailMaxElog && T.npart >= 0 && T.npart <= kPartCap && D.K >= 0 &&
      D.K <= kTailCap && D.npart >= 0 && D.npart <= kPartCap && D.elog >= 0 &&
      D.elog <= kTailMaxElog && (D.K == 0 || D.T1 + D.K == n) && B.fidx.size() == B.fw.size() &&
      B.bidx.size() == B.bw.size() &&
      (B.fptr.empty() || (int64_t)B.fptr.back() == (int64_t)B.fidx.size()) &&
      (B.bptr.empty() || (int64_t)B.bptr.back() == (int64_t)B.bidx.size());
  if (!shapes) {
    err = "load_plan: inconsistent array shapes in " + path;
    return false;
  }
  const bool bounds =
      all_within(Q.perm, 0, n) && all_within(Q.iperm, 0, n) && all_within(Q.Lidx, 0, n) &&
      all_within(Q.Fidx, 0, n) && all_within(Q.Bidx, 0, n) && all_within(Q.lvl, 0, n + 1) &&
      all_within(T.Eidx, 0, std::max<int64_t>(T.T0, 1)) && all_within(B.fidx, 0, 2 * n) &&
      all_within(B.bidx, 0, 2 * n) && all_within(B.lev, 0, Q.nlev + 1) &&
      program_in_bounds(T.fsteps, T.fslots, T.NT + T.npart) &&
      program_in_bounds(T.bsteps, T.bslots, T.NT + T.npart) &&
      program_in_bounds(D.fsteps, D.fslots, D.K + D.npart) &&
      program_in_bounds(D.bsteps, D.bslots, D.K + D.npart);
  if (!bounds) {
    err = "load_plan: an index out of range in " + path;
    return false;
  }
  P = std::move(Q);
  return true;
}

}  // namespace lapgpu
