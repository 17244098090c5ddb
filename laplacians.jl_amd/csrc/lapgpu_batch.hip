// lapgpu_batch.hip — device-resident graph store and batched solves, one
// workgroup per (graph, right-hand side) request (kernels: lapgpu_wg.hpp).
//
// A batch holds any number of uploaded graphs (la + LDLinv, level-ordered,
// with the whole LDLinv apply as one slot program).  lapgpu_batch_solve runs
// a set of requests of uneven sizes in ONE launch: many small graphs, many
// right-hand sides of one graph, or both; requests are ordered largest
// first so the long ones start early.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "lapgpu.h"
#include "lapgpu_lock.hpp"
#include "lapgpu_mem.hpp"
#include "lapgpu_plan.hpp"
#include "lapgpu_wg.hpp"

using namespace lapgpu;

// ---- scratch of lapgpu_batch_solve_device_noalloc -------------------------
// One array per device, holding the requests' graph descriptors, the request
// table and the per-request work vectors (8 n doubles each, plus up to 1024
// partial slots) of one call.  In
// host code `scratch` names a host-side stand-in: kernels and copies use the
// device address that hipGetSymbolAddress returns for the current device,
// looked up once per device when a batch is created there (scratch_setup).
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct GraphStore {
  int n = 0, nf = 0, nb = 0;
  int npart = 0;  // partial slots of the programs (work per request: 8 n + npart)
  int64_t bytes = 0;
  int *perm = nullptr, *Lptr = nullptr, *Lidx = nullptr;
  double *Lval = nullptr, *s = nullptr;
  TailStepD *fsteps = nullptr, *bsteps = nullptr;
  TailSlotD *fslots = nullptr, *bslots = nullptr;
  bool live = false;
};

// work vectors of one request: b, r, x, bestx, p, q, z and y (with the
// programs' partial slots past y[n))
size_t work_doubles(const GraphStore& G) { return 8 * (size_t)G.n + (size_t)G.npart; }

std::string g_batch_err;
constexpr int kBatchSlots = 4;

constexpr int64_t kNoAllocMaxReq = 1024;

std::mutex g_scratch_mu;           // guards g_scratch_ptr
std::vector<void*> g_scratch_ptr;  // per device: device address of scratch
// host staging of one call (static: the entry point allocates nothing);
// the calls hold g_noalloc_mu from start to end
std::mutex g_noalloc_mu;
WgGraphD g_na_graphs[kNoAllocMaxReq];
WgReqD g_na_reqs[kNoAllocMaxReq];
int64_t g_na_order[kNoAllocMaxReq];

// device address of this device's scratch (call with the device current)
hipError_t scratch_setup(int device, void** out) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if ((size_t)device < g_scratch_ptr.size() && g_scratch_ptr[device]) {
    *out = g_scratch_ptr[device];
    return hipSuccess;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize(device + 1, nullptr);
  g_scratch_ptr[device] = p;
  *out = p;
  return hipSuccess;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

// One launch slot: its own stream, request table, workspace and copy of the
// graph descriptor table, so that launches on different slots overlap on the
// device and adding or removing a graph never waits for a running launch.
struct BatchSlot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  WgGraphD* d_graphs = nullptr;
  size_t d_graphs_cap = 0;
  std::vector<WgGraphD> h_graphs;  // source of the last table upload
  uint64_t version = 0;            // store version of the table on the device
  std::vector<int32_t> gids;       // graphs the running launch reads
  WgReqD* d_reqs = nullptr;
  size_t reqs_cap = 0;
  WgReqD* h_reqs = nullptr;  // pinned
  size_t h_cap = 0;
  double* d_work = nullptr;
  size_t work_cap = 0;
  std::vector<int64_t> order;  // request table position -> request index
  std::vector<int64_t> pos;    // request index -> request table position
  // cancel flags, one per request table position, in coherent (uncached)
  // pinned host memory that the running kernel reads; earlier, smaller
  // arrays are kept until the batch is destroyed
  int* h_cancel = nullptr;
  const int* d_cancel = nullptr;
  size_t cancel_cap = 0;
  std::vector<int*> old_cancel;
  int64_t nreq = 0;
  bool busy = false;
  // host-buffer launches: device io and pinned staging
  double* d_io = nullptr;
  size_t io_cap = 0;
  double* h_io = nullptr;  // pinned
  size_t hio_cap = 0;
  size_t io_len = 0;       // doubles of x to hand back (0: device launch)
};

struct lapgpu_batch {
  // every entry point holds it, except add_graph while it plans on the host
  // (so graphs can be planned on other threads while launches go on)
  std::recursive_mutex mu;
  int device = 0;
  hipStream_t stream = nullptr;  // uploads and the synchronous entry points
  double wallclock_khz = 100000.0;
  std::string err;
  std::vector<GraphStore> graphs;
  uint64_t version = 1;  // bumped by every add and remove
  std::vector<BatchSlot> slots;
  double* d_io = nullptr;
  size_t d_io_cap = 0;
  void* scratch = nullptr;  // this device's scratch (lapgpu_batch_solve_device_noalloc)
};

namespace {

int berr(lapgpu_batch* B, int code, const std::string& m) {
  if (B)
    B->err = m;
  else
    g_batch_err = m;
  return code;
}

// the batch's lock for the rest of the calling entry point (none for a null B)
#define BATCH_LOCK(B)                                      \
  std::unique_lock<std::recursive_mutex> batch_lk_;        \
  if (B) batch_lk_ = std::unique_lock<std::recursive_mutex>((B)->mu)

#define BHIP(B, call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return berr(B, -10, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

// Device memory is stream-ordered (hipMallocAsync / hipFreeAsync): a free
// waits for the work on its stream, never for the whole device.  Graph
// uploads go through the batch's own stream, which add_graph synchronises
// before the graph can be named by a launch.
template <typename T>
int up(lapgpu_batch* B, T** dst, const T* src, size_t count, int64_t& bytes) {
  const size_t nb = sizeof(T) * (count ? count : 1);
  BHIP(B, hipMallocAsync((void**)dst, nb, B->stream));
  if (count)
    BHIP(B, hipMemcpyAsync(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice, B->stream));
  bytes += (int64_t)nb;
  return 0;
}

template <typename T>
int grow(lapgpu_batch* B, hipStream_t stream, T** p, size_t* cap, size_t want) {
  if (want <= *cap) return 0;
  if (*p) BHIP(B, hipFreeAsync(*p, stream));
  *p = nullptr;
  size_t c = std::max(want, *cap * 2);
  BHIP(B, hipMallocAsync((void**)p, sizeof(T) * c, stream));
  *cap = c;
  return 0;
}

// Pinned staging of at least `count` elements (kept while large enough).
template <typename T>
int grow_pinned(lapgpu_batch* B, T** p, size_t* cap, size_t count) {
  if (count <= *cap) return 0;
  lapgpu::pinned_put(*p, sizeof(T) * *cap);
  *p = nullptr;
  *cap = 0;
  size_t got = 0;
  BHIP(B, lapgpu::pinned_get((void**)p, sizeof(T) * count, &got));
  *cap = got / sizeof(T);
  return 0;
}

// A slot's cancel flags for at least `count` requests (the slot is idle).
int grow_cancel(lapgpu_batch* B, BatchSlot& S, size_t count) {
  if (count <= S.cancel_cap) return 0;
  const size_t c = std::max<size_t>(count, std::max<size_t>(1024, 2 * S.cancel_cap));
  int* p = nullptr;
  {
    LAPGPU_LOCK();
    BHIP(B, hipHostMalloc((void**)&p, sizeof(int) * c, hipHostMallocMapped | hipHostMallocCoherent));
  }
  void* d = nullptr;
  hipError_t e = hipHostGetDevicePointer(&d, p, 0);
  if (e != hipSuccess) {
    S.old_cancel.push_back(p);
    return berr(B, -10, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
  }
  if (S.h_cancel) S.old_cancel.push_back(S.h_cancel);
  S.h_cancel = p;
  S.d_cancel = (const int*)d;
  S.cancel_cap = c;
  return 0;
}

// Frees are ordered on `stream` (which must already wait for every launch
// that reads the graph).
void free_store(GraphStore& G, hipStream_t stream) {
  void* ps[] = {G.perm, G.Lptr, G.Lidx, G.Lval, G.s, G.fsteps, G.bsteps, G.fslots, G.bslots};
  for (void* p : ps)
    if (p) (void)hipFreeAsync(p, stream);
  G = GraphStore();
}

}  // namespace

extern "C" {

int lapgpu_batch_create(lapgpu_batch** out, int device) {
  if (!out) return berr(nullptr, -1, "lapgpu_batch_create: null output");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return berr(nullptr, -10, "lapgpu_batch_create: no HIP device visible");
  if (device < 0 || device >= count)
    return berr(nullptr, -1, "lapgpu_batch_create: device index out of range");
  LAPGPU_LOCK();
  lapgpu_batch* B = new lapgpu_batch();
  B->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = lapgpu::pool_setup(device);
  if (e == hipSuccess) e = scratch_setup(device, &B->scratch);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_batch_err = std::string("lapgpu_batch_create: ") + hipGetErrorString(e);
    delete B;
    return -10;
  }
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
    B->wallclock_khz = khz;
  B->slots.resize(kBatchSlots);
  for (BatchSlot& S : B->slots) {
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&S.done, hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    g_batch_err = std::string("lapgpu_batch_create: ") + hipGetErrorString(e);
    lapgpu_batch_destroy(B);
    return -10;
  }
  *out = B;
  return 0;
}

void lapgpu_batch_destroy(lapgpu_batch* B) {
  if (!B) return;
  LAPGPU_LOCK();
  (void)hipSetDevice(B->device);
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  for (BatchSlot& S : B->slots) {
    if (S.stream) {
      (void)hipStreamSynchronize(S.stream);
      void* ps[] = {S.d_graphs, S.d_reqs, S.d_work, S.d_io};
      for (void* p : ps)
        if (p) (void)hipFreeAsync(p, S.stream);
      (void)hipStreamSynchronize(S.stream);
    }
    lapgpu::pinned_put(S.h_reqs, sizeof(WgReqD) * S.h_cap);
    lapgpu::pinned_put(S.h_io, sizeof(double) * S.hio_cap);
    if (S.h_cancel) S.old_cancel.push_back(S.h_cancel);
    for (int* p : S.old_cancel) (void)hipHostFree(p);
    S.old_cancel.clear();
    S.h_cancel = nullptr;
    if (S.done) (void)hipEventDestroy(S.done);
    if (S.stream) (void)hipStreamDestroy(S.stream);
  }
  if (B->stream) {
    for (GraphStore& G : B->graphs) free_store(G, B->stream);
    if (B->d_io) (void)hipFreeAsync(B->d_io, B->stream);
    (void)hipStreamSynchronize(B->stream);
    (void)hipStreamDestroy(B->stream);
  }
  delete B;
}

const char* lapgpu_batch_last_error(lapgpu_batch* B) {
  return B ? B->err.c_str() : g_batch_err.c_str();
}

int lapgpu_batch_add_graph(lapgpu_batch* B, int64_t n, const int64_t* colptr,
                           const int64_t* rowval, const double* nzval,
                           int64_t nnzF, const int64_t* col,
                           const int64_t* fcolptr, const int64_t* frowval,
                           const double* fval, const double* d, int index_base,
                           int32_t* graph_id) {
  if (!B || !graph_id) return berr(B, -1, "lapgpu_batch_add_graph: null argument");
  std::string err;
  HostCsr L;
  Plan P;
  if (!csc_to_csr(n, colptr, rowval, nzval, index_base, L, err) ||
      !build_plan(L, nnzF, col, fcolptr, frowval, fval, d, index_base, 0, P, err)) {
    std::lock_guard<std::recursive_mutex> lk(B->mu);
    return berr(B, -1, err);
  }
  const bool planned = build_tail(P, INT64_MAX, INT64_MAX, kWgElog, kWgThreads) && P.tail.T0 == 0;
  std::lock_guard<std::recursive_mutex> lk(B->mu);
  if (!planned) return berr(B, -1, "lapgpu_batch_add_graph: could not build the slot program");
  GraphStore G;
  G.n = (int)n;
  G.nf = (int)P.tail.fsteps.size();
  G.nb = (int)P.tail.bsteps.size();
  G.npart = (int)P.tail.npart;
  static_assert(sizeof(TailStepD) == sizeof(TailStep), "step layout");
  static_assert(sizeof(TailSlotD) == sizeof(TailSlot), "slot layout");
  BHIP(B, hipSetDevice(B->device));
  int rc = 0;
  if ((rc = up(B, &G.perm, P.perm.data(), P.perm.size(), G.bytes)) ||
      (rc = up(B, &G.Lptr, P.Lptr.data(), P.Lptr.size(), G.bytes)) ||
      (rc = up(B, &G.Lidx, P.Lidx.data(), P.Lidx.size(), G.bytes)) ||
      (rc = up(B, &G.Lval, P.Lval.data(), P.Lval.size(), G.bytes)) ||
      (rc = up(B, &G.s, P.s.data(), P.s.size(), G.bytes)) ||
      (rc = up(B, &G.fsteps, (const TailStepD*)P.tail.fsteps.data(), P.tail.fsteps.size(), G.bytes)) ||
      (rc = up(B, &G.bsteps, (const TailStepD*)P.tail.bsteps.data(), P.tail.bsteps.size(), G.bytes)) ||
      (rc = up(B, &G.fslots, (const TailSlotD*)P.tail.fslots.data(), P.tail.fslots.size(), G.bytes)) ||
      (rc = up(B, &G.bslots, (const TailSlotD*)P.tail.bslots.data(), P.tail.bslots.size(), G.bytes))) {
    (void)hipStreamSynchronize(B->stream);
    free_store(G, B->stream);
    return rc;
  }
  // the plan's host vectors are the copies' sources
  BHIP(B, hipStreamSynchronize(B->stream));
  G.live = true;
  // reuse the id of a removed graph (launches that read it were ordered
  // before its frees; new launches get a new descriptor table)
  size_t id = 0;
  while (id < B->graphs.size() && B->graphs[id].live) ++id;
  if (id == B->graphs.size())
    B->graphs.push_back(G);
  else
    B->graphs[id] = G;
  ++B->version;
  *graph_id = (int32_t)id;
  B->err.clear();
  return 0;
}

int lapgpu_batch_remove_graph(lapgpu_batch* B, int32_t graph_id) {
  BATCH_LOCK(B);
  if (!B || graph_id < 0 || graph_id >= (int32_t)B->graphs.size() || !B->graphs[graph_id].live)
    return berr(B, -1, "lapgpu_batch_remove_graph: unknown graph");
  BHIP(B, hipSetDevice(B->device));
  // the frees wait (on the device) for the launches that read this graph
  for (BatchSlot& S : B->slots)
    if (S.busy && std::find(S.gids.begin(), S.gids.end(), graph_id) != S.gids.end())
      BHIP(B, hipStreamWaitEvent(B->stream, S.done, 0));
  free_store(B->graphs[graph_id], B->stream);
  ++B->version;
  return 0;
}

int lapgpu_batch_graph_info(lapgpu_batch* B, int32_t graph_id, int64_t* n,
                            int64_t* device_bytes) {
  BATCH_LOCK(B);
  if (!B || graph_id < 0 || graph_id >= (int32_t)B->graphs.size() || !B->graphs[graph_id].live)
    return berr(B, -1, "lapgpu_batch_graph_info: unknown graph");
  if (n) *n = B->graphs[graph_id].n;
  if (device_bytes) *device_bytes = B->graphs[graph_id].bytes;
  return 0;
}

}  // extern "C"

namespace {

int free_slot(lapgpu_batch* B) {
  for (int k = 0; k < (int)B->slots.size(); ++k)
    if (!B->slots[k].busy) return k;
  return -1;
}

// Enqueue nreq requests on slot `slot` (stream-ordered after whatever the
// caller already put on that stream).
int launch_on(lapgpu_batch* B, int slot, int64_t nreq, const int32_t* graph_ids,
              const double* const* d_b, double* const* d_x, double tol, int64_t maxits,
              double maxtime, int stag_test, int flags) {
  BatchSlot& S = B->slots[slot];
  const int ng = (int)B->graphs.size();
  int rc;
  if (S.version != B->version) {
    // this slot's own table (the slot is idle; other slots keep theirs);
    // h_graphs stays untouched until the slot's next launch, after collect
    S.h_graphs.resize(ng);
    for (int i = 0; i < ng; ++i) {
      const GraphStore& G = B->graphs[i];
      S.h_graphs[i] = WgGraphD{G.n, G.nf, G.nb, G.npart, G.perm, G.Lptr, G.Lidx, G.Lval, G.s,
                               G.fsteps, G.fslots, G.bsteps, G.bslots};
    }
    if ((rc = grow(B, S.stream, &S.d_graphs, &S.d_graphs_cap, (size_t)ng))) return rc;
    BHIP(B, hipMemcpyAsync(S.d_graphs, S.h_graphs.data(), sizeof(WgGraphD) * ng,
                           hipMemcpyHostToDevice, S.stream));
    S.version = B->version;
  }
  S.gids.assign(graph_ids, graph_ids + nreq);
  // largest first
  S.order.resize(nreq);
  std::iota(S.order.begin(), S.order.end(), 0);
  std::stable_sort(S.order.begin(), S.order.end(), [&](int64_t a, int64_t b) {
    return B->graphs[graph_ids[a]].n > B->graphs[graph_ids[b]].n;
  });
  size_t work = 0;
  for (int64_t i = 0; i < nreq; ++i) work += work_doubles(B->graphs[graph_ids[i]]);
  if ((rc = grow(B, S.stream, &S.d_work, &S.work_cap, work))) return rc;
  if ((rc = grow(B, S.stream, &S.d_reqs, &S.reqs_cap, (size_t)nreq))) return rc;
  if ((rc = grow_pinned(B, &S.h_reqs, &S.h_cap, (size_t)nreq))) return rc;
  if ((rc = grow_cancel(B, S, (size_t)nreq))) return rc;
  S.pos.resize(nreq);
  for (int64_t k = 0; k < nreq; ++k) {
    S.pos[S.order[k]] = k;
    __atomic_store_n(&S.h_cancel[k], 0, __ATOMIC_RELAXED);
  }
  // the cleared flags are visible before the launch that reads them
  std::atomic_thread_fence(std::memory_order_seq_cst);
  const bool has_mt = maxtime > 0 && std::isfinite(maxtime);
  size_t off = 0;
  for (int64_t k = 0; k < nreq; ++k) {
    const int64_t i = S.order[k];
    WgReqD R{};
    R.b = d_b[i];
    R.x = d_x[i];
    R.work = S.d_work + off;
    off += work_doubles(B->graphs[graph_ids[i]]);
    R.tol = tol;
    R.maxits = maxits;
    R.maxticks = has_mt ? (unsigned long long)(maxtime * B->wallclock_khz * 1e3) : 0ull;
    R.graph = graph_ids[i];
    R.stag_test = stag_test;
    R.center = (flags & LAPGPU_NO_CENTER) ? 0 : 1;
    R.has_maxtime = has_mt ? 1 : 0;
    R.cancel = S.d_cancel + k;
    S.h_reqs[k] = R;
  }
  BHIP(B, hipMemcpyAsync(S.d_reqs, S.h_reqs, sizeof(WgReqD) * nreq, hipMemcpyHostToDevice,
                         S.stream));
  hipLaunchKernelGGL(k_wg_pcg, dim3((unsigned)nreq), dim3(kWgThreads), 0, S.stream,
                     S.d_graphs, S.d_reqs);
  BHIP(B, hipGetLastError());
  BHIP(B, hipMemcpyAsync(S.h_reqs, S.d_reqs, sizeof(WgReqD) * nreq, hipMemcpyDeviceToHost,
                         S.stream));
  BHIP(B, hipEventRecord(S.done, S.stream));
  S.nreq = nreq;
  S.busy = true;
  return 0;
}

int check_ids(lapgpu_batch* B, int64_t nreq, const int32_t* graph_ids) {
  const int ng = (int)B->graphs.size();
  for (int64_t i = 0; i < nreq; ++i)
    if (graph_ids[i] < 0 || graph_ids[i] >= ng || !B->graphs[graph_ids[i]].live)
      return berr(B, -1, "lapgpu_batch: request names an unknown graph");
  return 0;
}

}  // namespace

extern "C" {

int lapgpu_batch_submit(lapgpu_batch* B, int64_t nreq, const int32_t* graph_ids,
                        const double* const* d_b, double* const* d_x, double tol,
                        int64_t maxits, double maxtime, int stag_test, int flags,
                        int32_t* slot_out) {
  BATCH_LOCK(B);
  if (!B || nreq <= 0 || !graph_ids || !d_b || !d_x || !slot_out)
    return berr(B, -1, "lapgpu_batch_submit: bad arguments");
  BHIP(B, hipSetDevice(B->device));
  int rc;
  if ((rc = check_ids(B, nreq, graph_ids))) return rc;
  const int slot = free_slot(B);
  if (slot < 0) return LAPGPU_BUSY;
  B->slots[slot].io_len = 0;
  if ((rc = launch_on(B, slot, nreq, graph_ids, d_b, d_x, tol, maxits, maxtime, stag_test,
                      flags)))
    return rc;
  *slot_out = slot;
  B->err.clear();
  return 0;
}

int lapgpu_batch_submit_host(lapgpu_batch* B, int64_t nreq, const int32_t* graph_ids,
                             const double* b_concat, double tol, int64_t maxits,
                             double maxtime, int stag_test, int flags, int32_t* slot_out) {
  BATCH_LOCK(B);
  if (!B || nreq <= 0 || !graph_ids || !b_concat || !slot_out)
    return berr(B, -1, "lapgpu_batch_submit_host: bad arguments");
  BHIP(B, hipSetDevice(B->device));
  int rc;
  if ((rc = check_ids(B, nreq, graph_ids))) return rc;
  const int slot = free_slot(B);
  if (slot < 0) return LAPGPU_BUSY;
  BatchSlot& S = B->slots[slot];
  size_t total = 0;
  for (int64_t i = 0; i < nreq; ++i) total += (size_t)B->graphs[graph_ids[i]].n;
  if ((rc = grow(B, S.stream, &S.d_io, &S.io_cap, 2 * total))) return rc;
  if ((rc = grow_pinned(B, &S.h_io, &S.hio_cap, total))) return rc;
  std::memcpy(S.h_io, b_concat, sizeof(double) * total);
  BHIP(B, hipMemcpyAsync(S.d_io, S.h_io, sizeof(double) * total, hipMemcpyHostToDevice, S.stream));
  std::vector<const double*> pb(nreq);
  std::vector<double*> px(nreq);
  size_t off = 0;
  for (int64_t i = 0; i < nreq; ++i) {
    pb[i] = S.d_io + off;
    px[i] = S.d_io + total + off;
    off += (size_t)B->graphs[graph_ids[i]].n;
  }
  if ((rc = launch_on(B, slot, nreq, graph_ids, pb.data(), px.data(), tol, maxits, maxtime,
                      stag_test, flags)))
    return rc;
  // x comes back into the pinned staging after the kernel
  BHIP(B, hipMemcpyAsync(S.h_io, S.d_io + total, sizeof(double) * total, hipMemcpyDeviceToHost,
                         S.stream));
  BHIP(B, hipEventRecord(S.done, S.stream));
  S.io_len = total;
  *slot_out = slot;
  B->err.clear();
  return 0;
}

int lapgpu_batch_cancel(lapgpu_batch* B, int32_t slot, int64_t index) {
  BATCH_LOCK(B);
  if (!B || slot < 0 || slot >= (int32_t)B->slots.size())
    return berr(B, -1, "lapgpu_batch_cancel: bad slot");
  BatchSlot& S = B->slots[slot];
  if (!S.busy) return berr(B, -1, "lapgpu_batch_cancel: slot has no launch");
  if (index < 0 || index >= S.nreq) return berr(B, -1, "lapgpu_batch_cancel: request index out of range");
  __atomic_store_n(&S.h_cancel[S.pos[index]], 1, __ATOMIC_RELEASE);
  B->err.clear();
  return 0;
}

int lapgpu_batch_query(lapgpu_batch* B, int32_t slot, int* done) {
  BATCH_LOCK(B);
  if (!B || slot < 0 || slot >= (int32_t)B->slots.size() || !done)
    return berr(B, -1, "lapgpu_batch_query: bad arguments");
  BatchSlot& S = B->slots[slot];
  if (!S.busy) return berr(B, -1, "lapgpu_batch_query: slot has no launch");
  hipError_t e = hipEventQuery(S.done);
  if (e == hipSuccess)
    *done = 1;
  else if (e == hipErrorNotReady)
    *done = 0;
  else
    return berr(B, -10, std::string("lapgpu_batch_query: ") + hipGetErrorString(e));
  return 0;
}

int lapgpu_batch_collect(lapgpu_batch* B, int32_t slot, double* x_concat, int64_t* its_out,
                         double* relres_out, int32_t* stop_out) {
  BATCH_LOCK(B);
  if (!B || slot < 0 || slot >= (int32_t)B->slots.size())
    return berr(B, -1, "lapgpu_batch_collect: bad arguments");
  BatchSlot& S = B->slots[slot];
  if (!S.busy) return berr(B, -1, "lapgpu_batch_collect: slot has no launch");
  BHIP(B, hipSetDevice(B->device));
  BHIP(B, hipEventSynchronize(S.done));
  if (S.io_len && x_concat) std::memcpy(x_concat, S.h_io, sizeof(double) * S.io_len);
  for (int64_t k = 0; k < S.nreq; ++k) {
    const int64_t i = S.order[k];
    if (its_out) its_out[i] = S.h_reqs[k].its;
    if (relres_out) relres_out[i] = S.h_reqs[k].nr;
    if (stop_out) stop_out[i] = S.h_reqs[k].stop;
  }
  S.busy = false;
  B->err.clear();
  return 0;
}

int lapgpu_batch_solve_device(lapgpu_batch* B, int64_t nreq,
                              const int32_t* graph_ids,
                              const double* const* d_b, double* const* d_x,
                              double tol, int64_t maxits, double maxtime,
                              int stag_test, int flags, int64_t* its_out,
                              double* relres_out, int32_t* stop_out) {
  BATCH_LOCK(B);
  if (!B || nreq < 0) return berr(B, -1, "lapgpu_batch_solve: bad arguments");
  if (nreq == 0) return 0;
  int32_t slot = -1;
  int rc;
  rc = lapgpu_batch_submit(B, nreq, graph_ids, d_b, d_x, tol, maxits, maxtime, stag_test,
                           flags, &slot);
  if (rc == LAPGPU_BUSY)
    return berr(B, -2, "lapgpu_batch_solve: all launch slots hold uncollected launches");
  if (rc) return rc;
  return lapgpu_batch_collect(B, slot, nullptr, its_out, relres_out, stop_out);
}

int lapgpu_batch_solve(lapgpu_batch* B, int64_t nreq, const int32_t* graph_ids,
                       const double* b_concat, double* x_concat, double tol,
                       int64_t maxits, double maxtime, int stag_test, int flags,
                       int64_t* its_out, double* relres_out, int32_t* stop_out) {
  BATCH_LOCK(B);
  if (!B || nreq < 0 || (nreq && (!graph_ids || !b_concat || !x_concat)))
    return berr(B, -1, "lapgpu_batch_solve: bad arguments");
  if (nreq == 0) return 0;
  BHIP(B, hipSetDevice(B->device));
  const int ng = (int)B->graphs.size();
  size_t total = 0;
  for (int64_t i = 0; i < nreq; ++i) {
    if (graph_ids[i] < 0 || graph_ids[i] >= ng || !B->graphs[graph_ids[i]].live)
      return berr(B, -1, "lapgpu_batch_solve: request names an unknown graph");
    total += (size_t)B->graphs[graph_ids[i]].n;
  }
  int rc;
  if ((rc = grow(B, B->stream, &B->d_io, &B->d_io_cap, 2 * total))) return rc;
  BHIP(B, hipMemcpyAsync(B->d_io, b_concat, sizeof(double) * total, hipMemcpyHostToDevice,
                         B->stream));
  BHIP(B, hipStreamSynchronize(B->stream));
  std::vector<const double*> pb(nreq);
  std::vector<double*> px(nreq);
  size_t off = 0;
  for (int64_t i = 0; i < nreq; ++i) {
    pb[i] = B->d_io + off;
    px[i] = B->d_io + total + off;
    off += (size_t)B->graphs[graph_ids[i]].n;
  }
  rc = lapgpu_batch_solve_device(B, nreq, graph_ids, pb.data(), px.data(), tol, maxits,
                                 maxtime, stag_test, flags, its_out, relres_out, stop_out);
  if (rc) return rc;
  BHIP(B, hipMemcpyAsync(x_concat, B->d_io + total, sizeof(double) * total,
                         hipMemcpyDeviceToHost, B->stream));
  BHIP(B, hipStreamSynchronize(B->stream));
  return 0;
}


/* Allocation-free form of lapgpu_batch_solve_device: the same kernel and
 * arithmetic, with its work vectors, request table and graph descriptors
 * in the device's scratch array (kScratchBytes) and host staging in static
 * arrays; no hipMalloc, hipFree or pinned allocation. */
int lapgpu_batch_solve_device_noalloc(lapgpu_batch* B, int64_t nreq, const int32_t* graph_ids,
                                      const double* const* d_b, double* const* d_x, double tol,
                                      int64_t maxits, double maxtime, int stag_test, int flags,
                                      int64_t* its_out, double* relres_out, int32_t* stop_out) {
  BATCH_LOCK(B);
  if (!B || nreq < 0 || (nreq && (!graph_ids || !d_b || !d_x)))
    return berr(B, -1, "lapgpu_batch_solve_device_noalloc: bad arguments");
  if (nreq == 0) return 0;
  if (nreq > kNoAllocMaxReq)
    return berr(B, -1, "lapgpu_batch_solve_device_noalloc: more than 1024 requests in one call");
  int rc;
  if ((rc = check_ids(B, nreq, graph_ids))) return rc;
  // scratch layout: descriptors | requests | work vectors (8 n + npart
  // doubles each)
  const size_t off_reqs = align256(sizeof(WgGraphD) * (size_t)nreq);
  const size_t off_work = off_reqs + align256(sizeof(WgReqD) * (size_t)nreq);
  size_t need = off_work;
  for (int64_t i = 0; i < nreq; ++i) need += sizeof(double) * work_doubles(B->graphs[graph_ids[i]]);
  if (need > kScratchBytes)
    return berr(B, -1, "lapgpu_batch_solve_device_noalloc: the requests need " + std::to_string(need) +
                           " bytes of scratch, more than its " + std::to_string(kScratchBytes) +
                           " (at most about 1,000,000 vertices over all requests of a call)");
  if (!B->scratch) return berr(B, -2, "lapgpu_batch_solve_device_noalloc: no scratch on this device");
  BHIP(B, hipSetDevice(B->device));
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  unsigned char* base = static_cast<unsigned char*>(B->scratch);
  WgGraphD* d_graphs = reinterpret_cast<WgGraphD*>(base);
  WgReqD* d_reqs = reinterpret_cast<WgReqD*>(base + off_reqs);
  double* work = reinterpret_cast<double*>(base + off_work);
  // largest first, as the other entry points order them (an in-place sort
  // with the request index as tie-break: the same order as a stable sort)
  for (int64_t i = 0; i < nreq; ++i) g_na_order[i] = i;
  std::sort(g_na_order, g_na_order + nreq, [&](int64_t a, int64_t b) {
    const int na = B->graphs[graph_ids[a]].n, nb = B->graphs[graph_ids[b]].n;
    return na != nb ? na > nb : a < b;
  });
  const bool has_mt = maxtime > 0 && std::isfinite(maxtime);
  size_t off = 0;
  for (int64_t k = 0; k < nreq; ++k) {
    const int64_t i = g_na_order[k];
    const GraphStore& G = B->graphs[graph_ids[i]];
    // one descriptor per request (no table of every graph to upload)
    g_na_graphs[k] = WgGraphD{G.n, G.nf, G.nb, G.npart, G.perm, G.Lptr, G.Lidx, G.Lval, G.s,
                              G.fsteps, G.fslots, G.bsteps, G.bslots};
    WgReqD R{};
    R.b = d_b[i];
    R.x = d_x[i];
    R.work = work + off;
    off += work_doubles(G);
    R.tol = tol;
    R.maxits = maxits;
    R.maxticks = has_mt ? (unsigned long long)(maxtime * B->wallclock_khz * 1e3) : 0ull;
    R.graph = (int)k;
    R.stag_test = stag_test;
    R.center = (flags & LAPGPU_NO_CENTER) ? 0 : 1;
    R.has_maxtime = has_mt ? 1 : 0;
    g_na_reqs[k] = R;
  }
  BHIP(B, hipMemcpyAsync(d_graphs, g_na_graphs, sizeof(WgGraphD) * nreq, hipMemcpyHostToDevice,
                         B->stream));
  BHIP(B, hipMemcpyAsync(d_reqs, g_na_reqs, sizeof(WgReqD) * nreq, hipMemcpyHostToDevice,
                         B->stream));
  hipLaunchKernelGGL(k_wg_pcg, dim3((unsigned)nreq), dim3(kWgThreads), 0, B->stream, d_graphs,
                     d_reqs);
  BHIP(B, hipGetLastError());
  BHIP(B, hipMemcpyAsync(g_na_reqs, d_reqs, sizeof(WgReqD) * nreq, hipMemcpyDeviceToHost,
                         B->stream));
  BHIP(B, hipStreamSynchronize(B->stream));
  for (int64_t k = 0; k < nreq; ++k) {
    const int64_t i = g_na_order[k];
    if (its_out) its_out[i] = g_na_reqs[k].its;
    if (relres_out) relres_out[i] = g_na_reqs[k].nr;
    if (stop_out) stop_out[i] = g_na_reqs[k].stop;
  }
  B->err.clear();
  return 0;
}

}  // extern "C"
