// lapgpu.hip — C-ABI host side of the gfx950 solve phase of approxchol_lap.
//
// One context = one HIP stream (default priority) + one uploaded graph
// (la and its LDLinv in level order) + the PCG workspace.  lapgpu_solve runs
// the whole of pcg (src/pcg.jl:256-445) on the device: the host only
// enqueues kernels and polls a stop flag one iteration behind, so the GPU
// never waits for the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "lapgpu.h"
#include "lapgpu_chain.hpp"
#include "lapgpu_lock.hpp"
#include "lapgpu_mem.hpp"
#include "lapgpu_multi.hpp"
#include "lapgpu_plan.hpp"

using namespace lapgpu;

static_assert(kTailThreads == kTailLanes, "tail workgroup width = lanes of a tail step");
static_assert(kTailCapLds == kTailCap, "tail capacity: planner and kernels agree");
static_assert(kPartSlots == kPartCap, "partial slots: planner and kernels agree");

namespace lapgpu {

hipError_t pool_setup(int device) {
  static std::mutex mu;
  static std::vector<bool> done;
  std::lock_guard<std::mutex> g(mu);
  if ((size_t)device < done.size() && done[device]) return hipSuccess;
  hipMemPool_t pool;
  hipError_t e = hipDeviceGetDefaultMemPool(&pool, device);
  if (e != hipSuccess) return e;
  uint64_t keep = UINT64_MAX;
  if ((e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep)) != hipSuccess)
    return e;
  if ((size_t)device >= done.size()) done.resize(device + 1, false);
  done[device] = true;
  return hipSuccess;
}

namespace {
std::mutex g_pinned_mu;
std::map<size_t, std::vector<void*>> g_pinned;  // size class -> free blocks

size_t pinned_class(size_t bytes) {
  size_t c = 256;
  while (c < bytes) c <<= 1;
  return c;
}
}  // namespace

hipError_t pinned_get(void** p, size_t bytes, size_t* cap) {
  const size_t c = pinned_class(bytes);
  if (cap) *cap = c;
  {
    std::lock_guard<std::mutex> g(g_pinned_mu);
    auto it = g_pinned.find(c);
    if (it != g_pinned.end() && !it->second.empty()) {
      *p = it->second.back();
      it->second.pop_back();
      return hipSuccess;
    }
  }
  LAPGPU_LOCK();
  return hipHostMalloc(p, c, hipHostMallocDefault);
}

void pinned_put(void* p, size_t bytes) {
  if (!p) return;
  std::lock_guard<std::mutex> g(g_pinned_mu);
  g_pinned[pinned_class(bytes)].push_back(p);
}

}  // namespace lapgpu

namespace {

std::string g_create_err;
constexpr int kMultiMax = 64;  // columns of one multi-RHS chunk

template <typename T>
struct DevArr {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace

struct lapgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  double wallclock_khz = 100000.0;
  std::string err;

  HostCsr L;
  bool has_L = false, has_F = false;
  Plan plan;
  int64_t n = 0;
  int64_t max_tail_steps = 24;
  int G_spmv = 8, G_ldl = 8;
  int precision = LAPGPU_FP64, replace_every = 20;
  // FP32 mode: relative residual below which the SpMV is FP64.  The FP32
  // SpMV's rounding costs CG iterations once the residual is small: on a
  // 2-D grid with 64 right-hand sides the slowest column needed 62
  // iterations at 1e-4 against 47 at 1e-2 (FP64: 45)
  double switch_tol = 1e-2;

  // cancellation (lapgpu_cancel, from another thread): one flag per column
  // of the running solve call; cancel_any is polled by the solve loop
  std::mutex cancel_mu;
  std::atomic<bool> cancel_any{false};
  std::vector<char> cancel_cols;
  bool running = false;
  // global column index of column 0 of the multi-RHS chunk being solved
  int64_t col_base = 0, chunk0 = 0;
  // FP64 true relative residual of each column the last solve call returned
  std::vector<double> last_true;

  // device graph
  int *perm = nullptr, *Lptr = nullptr, *Lidx = nullptr;
  double* Lval = nullptr;
  int* lvl = nullptr;
  double* s = nullptr;
  // level groups before the tail (composite rows; see LevelBlocks)
  struct GroupLaunch {
    int lo, hi;
    int Gf, nhf, hoff_f;
    int Gb, nhb, hoff_b;
    // multi-RHS: entry slices per row (before the 64/K bound) and the rows
    // longer than kMHeavyLen, run one workgroup per row
    int Sf, mnhf, mhoff_f;
    int Sb, mnhb, mhoff_b;
  };
  std::vector<GroupLaunch> groups;
  int *gFptr = nullptr, *gFidx = nullptr, *gBptr = nullptr, *gBidx = nullptr,
      *heavy = nullptr, *mheavy = nullptr;
  double *gFw = nullptr, *gBw = nullptr;
  float *gFw32 = nullptr, *gBw32 = nullptr, *z32 = nullptr;
  double* zF = nullptr;  // backward output yF (double path)
  // single-RHS sweeps: runs of narrow level groups in one launch each
  // (lapgpu_chain.hpp); off (env LAPGPU_CHAIN=0, or after a wait ran out):
  // one launch per group.  A group is narrow when it has at most
  // chain_narrow rows; a run is two or more consecutive narrow groups.
  // group_run: each group's run (-1: a launch of its own); runs[r]: its
  // first and last group; tkF/tkB: its tickets in either sweep; ctl: the
  // timeout flag, the poll limit of a wait (chain_poll_max) and the runs'
  // ticket counters (forward run r at 2 + r, backward at 2 + nruns + r);
  // chain_lo/hi: the positions the runs write.
  // (Multi-RHS sweeps keep one launch per group: chained runs measured
  // slower there at every threshold.)
  bool use_chain = true;
  int chain_narrow = 1024;
  // shorter runs stay one launch per group: measured on 1e6..2e6-vertex
  // expanders and 3-D grids, runs of 44 or more groups were 6-27 % faster,
  // runs of 14-34 groups 2-4 % slower
  int chain_min_run = 40;
  int chain_poll_max = kChainPollMax;
  // multi-RHS level groups: heavy and light rows in one launch for K up to
  // this many columns (env LAPGPU_MGROUP_MERGE_K).  Measured (1024-thread
  // workgroups against a 256-thread launch plus a heavy-row launch, 3.4M
  // expander and 1.3M 3-D grid, FP32-inner LDLinv apply): K = 8 -14/-17 %,
  // 16 -11/-15 %, 32 -6/-11 %, 64 +4/+11 %.
  int mgroup_merge_k = 32;
  std::vector<int> group_run;
  std::vector<std::pair<int, int>> runs;
  std::vector<std::pair<int, int>> tkF, tkB;  // ticket ranges per run
  ChainTicket *chainF = nullptr, *chainB = nullptr;
  int64_t chain_lo = 0, chain_hi = 0;
  int* chain_ctl = nullptr;
  int32_t* hchain = nullptr;
  // level-group limits: a level joins the group before it while the group
  // spans at most block_depth levels, its composite rows stay within
  // block_growth x its entries and add at most block_extra entries (each
  // launch costs ~4 us, each entry ~8 ps on MI355X: narrow levels join,
  // wide ones get launches of their own)
  int block_depth = 32;
  double block_growth = 16.0;
  int64_t block_extra = 196608;
  // tail program
  int *Eptr = nullptr, *Eidx = nullptr;
  double *Ew = nullptr, *ext = nullptr;
  TailStepD *fsteps = nullptr, *bsteps = nullptr;
  TailSlotD *fslots = nullptr, *bslots = nullptr;
  // the tail's dense block (P.dense): D (K rows, leading dimension ldd) and
  // the block's inputs (ldd entries, zero beyond K); the largest block the
  // planner may choose (0: none)
  int64_t dense_max = kTailCap;
  int64_t ldd = 0;
  double *D = nullptr, *dvec = nullptr;
  float *D32 = nullptr, *dvec32 = nullptr;
  // FP32-inner mode copies (built by lapgpu_set_precision)
  bool f32_ready = false;
  float *Lval32 = nullptr, *s32 = nullptr, *Ew32 = nullptr, *y32 = nullptr, *ext32 = nullptr,
        *pf = nullptr;
  TailSlotF *fslots32 = nullptr, *bslots32 = nullptr;
  double* r0 = nullptr;
  // one PCG iteration captured as a HIP graph (rebuilt after each upload);
  // iter_exec_rep is the FP32-mode iteration with a residual replacement
  hipGraphExec_t iter_exec = nullptr, iter_exec_rep = nullptr;
  bool use_graph = true;

  // workspace
  double *b = nullptr, *r = nullptr, *x = nullptr, *bestx = nullptr,
         *p = nullptr, *q = nullptr, *y = nullptr, *z = nullptr,
         *io = nullptr, *part = nullptr;
  DevState* st = nullptr;
  DevState* hst = nullptr;        // pinned host mirror
  int32_t* hflags = nullptr;      // pinned ring of stop flags
  static constexpr int kRing = 4;
  hipEvent_t ring_ev[kRing] = {};
  // waits on the device: spin (false) or sleep on blocking-sync events
  bool blocking = false;
  hipEvent_t sync_ev = nullptr;

  // multi-RHS workspace (lapgpu_solve_multi): K interleaved right-hand
  // sides of length n; the tail CSRs are built on first use
  struct Multi {
    int64_t cap = 0;  // elements of each vector (n * K)
    int exec_K = 0;   // K of the captured iterations
    double *r = nullptr, *x = nullptr, *bestx = nullptr, *p = nullptr, *q = nullptr,
           *y = nullptr, *z = nullptr, *io = nullptr, *part = nullptr;
    DevState* st = nullptr;
    DevState* hst = nullptr;  // pinned, 64 states
    int* flags = nullptr;
    double* ext = nullptr;  // tail inputs, NT * K
    // FP32-inner mode: sweep buffers, the FP32 copy of p, r0 = b - mean(b)
    int64_t cap32 = 0;
    float *y32 = nullptr, *z32 = nullptr, *pf = nullptr, *ext32 = nullptr;
    double* r0 = nullptr;
    // dense block: its inputs (ldd x K, zero beyond row K) and the split
    // partial products
    double *dvec = nullptr, *dpart = nullptr;
    float *dvec32 = nullptr, *dpart32 = nullptr;
    // captured iteration (exec[1]: FP32 mode with a residual replacement)
    hipGraphExec_t exec[2] = {nullptr, nullptr};
  } m;

  // profiling (LAPGPU_PROFILE): events around the kernel classes of one
  // iteration; ev[5..7] split the sweeps into forward groups, tail inputs,
  // tail and backward groups, ev[8..9] the tail into its sparse forward, the
  // dense block's product and its sparse backward
  hipEvent_t ev[10] = {};
  bool prof_sweeps = false;
  double t_spmv = 0, t_ldl = 0, t_vec = 0, t_total = 0;
  double t_fgroups = 0, t_tail_ext = 0, t_tail = 0, t_bgroups = 0;
  double t_tail_fwd = 0, t_dense = 0, t_tail_bwd = 0;
  int64_t t_its = 0;
};

namespace {

int set_err(lapgpu_ctx* c, int code, const std::string& msg) {
  if (c)
    c->err = msg;
  else
    g_create_err = msg;
  return code;
}

#define HIPC(ctx, call)                                                      \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return set_err(ctx, -10, std::string(#call) + ": " +                   \
                                   hipGetErrorString(e_));                   \
  } while (0)

// Wait for an event: HIP's own wait, or, in the blocking wait mode, queries
// with 20 us sleeps in between (the thread leaves its core to other work;
// a wait ends at most ~20 us after the event).
hipError_t wait_event(lapgpu_ctx* c, hipEvent_t ev) {
  if (!c->blocking) return hipEventSynchronize(ev);
  for (;;) {
    const hipError_t e = hipEventQuery(ev);
    if (e != hipErrorNotReady) return e;
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

// Wait for everything queued on the context's stream.
hipError_t sync_stream(lapgpu_ctx* c) {
  if (!c->blocking) return hipStreamSynchronize(c->stream);
  hipError_t e = hipEventRecord(c->sync_ev, c->stream);
  return e != hipSuccess ? e : wait_event(c, c->sync_ev);
}

// Device buffers are stream-ordered on the context's stream (a free waits
// for the context's own work only, never for the whole device).
void dev_free(lapgpu_ctx* c, void* p) {
  if (p) (void)hipFreeAsync(p, c->stream);
}

template <typename T>
int dev_upload(lapgpu_ctx* c, T** dst, const std::vector<T>& src) {
  dev_free(c, *dst);
  *dst = nullptr;
  size_t bytes = sizeof(T) * (src.empty() ? 1 : src.size());
  HIPC(c, hipMallocAsync((void**)dst, bytes, c->stream));
  if (!src.empty()) {
    HIPC(c, hipMemcpyAsync(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice,
                           c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));  // src is the caller's host buffer
  }
  return 0;
}

template <typename T>
int dev_alloc(lapgpu_ctx* c, T** dst, size_t count) {
  dev_free(c, *dst);
  *dst = nullptr;
  HIPC(c, hipMallocAsync((void**)dst, sizeof(T) * (count ? count : 1), c->stream));
  return 0;
}

// count elements, zero-filled (at upload time; never inside a capture)
template <typename T>
int dev_zeros(lapgpu_ctx* c, T** dst, size_t count) {
  int rc = dev_alloc(c, dst, count);
  if (rc) return rc;
  HIPC(c, hipMemsetAsync(*dst, 0, sizeof(T) * (count ? count : 1), c->stream));
  return 0;
}

// tail inputs: NT entries, and up to ldd (zero) when the dense block takes
// them as they are (no sparse part before it)
size_t ext_len(const lapgpu_ctx* c) {
  const TailProgram& TP = c->plan.tail;
  size_t n = (size_t)std::max<int64_t>(TP.NT, 1);
  if (c->plan.dense.K > 0 && TP.NT1 == 0) n = std::max<size_t>(n, (size_t)c->ldd);
  return n;
}

int pick_group(double avg) {
  // lanes per row: the power of two nearest to the mean row length
  int g = 2;
  while (g < 64 && (double)g * 1.5 < avg) g *= 2;
  if (g < 4) g = 4;
  return g;
}

int grid_for(int64_t items, int per_block) {
  int64_t b = (items + per_block - 1) / per_block;
  return (int)(b < 1 ? 1 : b);
}

// ---- templated launch helpers ----------------------------------------------

// Calls F with the compile-time lane count G_ = g (F may name G_).
#define LAPGPU_DISPATCH_G(g, F, ...)                   \
  switch (g) {                                         \
    case 4: { constexpr int G_ = 4; F(__VA_ARGS__); break; }    \
    case 8: { constexpr int G_ = 8; F(__VA_ARGS__); break; }    \
    case 16: { constexpr int G_ = 16; F(__VA_ARGS__); break; }  \
    case 32: { constexpr int G_ = 32; F(__VA_ARGS__); break; }  \
    default: { constexpr int G_ = 64; F(__VA_ARGS__); break; }  \
  }

// Calls F with the compile-time entries per lane E_ = 2^elog of a tail
// program (F may name E_).
#define LAPGPU_DISPATCH_E(elog, F, ...)                              \
  switch (elog) {                                                    \
    case 0: { constexpr int E_ = 1; F(__VA_ARGS__); break; }         \
    case 1: { constexpr int E_ = 2; F(__VA_ARGS__); break; }         \
    default: { constexpr int E_ = 4; F(__VA_ARGS__); break; }        \
  }

template <int G>
void launch_spmv(lapgpu_ctx* c, const double* p, double* q) {
  if (c->precision == LAPGPU_FP32_INNER)
    hipLaunchKernelGGL(k_spmv_dot_mixed<G>, dim3(kRedBlocks), dim3(kBlock), 0, c->stream,
                       (int)c->n, c->Lptr, c->Lidx, c->Lval32, c->pf, c->Lval, p, q, c->part,
                       c->st);
  else
    hipLaunchKernelGGL((k_spmv_dot<G, double>), dim3(kRedBlocks), dim3(kBlock), 0,
                       c->stream, (int)c->n, c->Lptr, c->Lidx, c->Lval, p, p, q,
                       c->part, c->st);
}

// q = la * p (p.q partials); FP32 mode reads the FP32 copies of la and p
void spmv(lapgpu_ctx* c, const double* p, double* q) {
  LAPGPU_DISPATCH_G(c->G_spmv, launch_spmv<G_>, c, p, q);
}

// the FP64 true residual of the returned x (bestx after the final copy)
template <int G>
void launch_true_resid(lapgpu_ctx* c) {
  hipLaunchKernelGGL(k_true_resid<G>, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, (int)c->n,
                     c->Lptr, c->Lidx, c->Lval, c->bestx, c->b, c->part, c->st);
  hipLaunchKernelGGL(k_scalar_true, dim3(1), dim3(kScalarThreads), 0, c->stream, c->part, c->st);
}

template <int G>
void launch_resid(lapgpu_ctx* c, int always) {
  hipLaunchKernelGGL(k_resid<G>, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, (int)c->n,
                     always, c->Lptr, c->Lidx, c->Lval, c->x, c->r0, c->r, c->part, c->st);
}

template <int G, typename T, typename TB>
void launch_group(lapgpu_ctx* c, int lo, int hi, const int* heavy, int nh, const int* ptr,
                  const int* idx, const T* w, const T* A, const TB* B, T* out) {
  const int64_t threads = (int64_t)nh * kWave + (int64_t)(hi - lo) * G;
  hipLaunchKernelGGL((k_group<G, T, TB>), dim3(grid_for(threads, kBlock)), dim3(kBlock), 0,
                     c->stream, (int)c->n, lo, hi, heavy, nh, ptr, idx, w, A, B, out);
}

// the dense block in precision T: D and its input vector
template <typename T>
struct DenseArgs {
  const T* D;
  T* dvec;
};
template <typename T>
DenseArgs<T> dense_args(lapgpu_ctx* c);
template <>
DenseArgs<double> dense_args<double>(lapgpu_ctx* c) { return {c->D, c->dvec}; }
template <>
DenseArgs<float> dense_args<float>(lapgpu_ctx* c) { return {c->D32, c->dvec32}; }

bool chain_on(const lapgpu_ctx* c) { return c->use_chain && !c->runs.empty() && c->chain_ctl; }

// the sentinel into the runs' output positions (elements [lo K, hi K) of
// y and z) and the runs' counters to zero, before a pair of sweeps
template <typename T>
void launch_chain_prep(lapgpu_ctx* c, int K, T* y, T* z) {
  const int64_t lo = c->chain_lo * K, hi = c->chain_hi * K;
  hipLaunchKernelGGL(k_chain_prep<T>, dim3((int)std::min<int64_t>(1024, (hi - lo + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, c->stream, lo, hi, y, z, c->chain_ctl,
                     (int)(2 * c->runs.size()));
}

template <typename T, typename TB>
void launch_chain(lapgpu_ctx* c, const ChainTicket* tk, std::pair<int, int> range, int ctr,
                  const int* ptr, const int* idx, const T* w, const T* A, const TB* B, T* out) {
  hipLaunchKernelGGL((k_chain<T, TB>), dim3(std::min(range.second - range.first, kChainGrid)),
                     dim3(kBlock), 0, c->stream, (int)c->n, tk, range.first, range.second,
                     c->chain_ctl, ctr, c->heavy, ptr, idx, w, A, B, out);
}

template <typename T, typename S>
void launch_sweeps(lapgpu_ctx* c, const double* rin, T* y, T* z, const T* gFw, const T* gBw,
                   const T* sv, const T* Ew, T* ext, const S* fslots, const S* bslots) {
  const Plan& P = c->plan;
  const bool chain = chain_on(c);
  const int nruns = (int)c->runs.size();
  if (chain) launch_chain_prep<T>(c, 1, y, z);
  for (int k = 0; k < (int)c->groups.size(); ++k) {
    const auto& g = c->groups[k];
    const int r = chain ? c->group_run[k] : -1;
    if (r >= 0) {
      if (k == c->runs[r].first)
        launch_chain<T, double>(c, c->chainF, c->tkF[r], r, c->gFptr, c->gFidx, gFw, y, rin, y);
      continue;
    }
    const int* hv = c->heavy + g.hoff_f;
    LAPGPU_DISPATCH_G(g.Gf, (launch_group<G_, T, double>), c, g.lo, g.hi, hv, g.nhf, c->gFptr,
                      c->gFidx, gFw, y, rin, y);
  }
  const TailProgram& TP = P.tail;
  if (c->prof_sweeps) (void)hipEventRecord(c->ev[5], c->stream);
  if (TP.NT > 0) {
    const int G = c->G_ldl;
    if (G <= 8)
      hipLaunchKernelGGL((k_tail_ext<8, T>), dim3(grid_for(TP.NT * 8, kBlock)), dim3(kBlock), 0,
                         c->stream, (int)TP.T0, (int)TP.NT, c->Eptr, c->Eidx, Ew, rin, y, ext);
    else
      hipLaunchKernelGGL((k_tail_ext<32, T>), dim3(grid_for(TP.NT * 32, kBlock)), dim3(kBlock),
                         0, c->stream, (int)TP.T0, (int)TP.NT, c->Eptr, c->Eidx, Ew, rin, y, ext);
    if (c->prof_sweeps) (void)hipEventRecord(c->ev[6], c->stream);
    const DenseBlock& DB = P.dense;
    if (DB.K == 0) {
      LAPGPU_DISPATCH_E(TP.elog, hipLaunchKernelGGL, (k_tail<T, S, E_>), dim3(1),
                        dim3(kTailThreads), 0, c->stream, (int)TP.T0, (int)TP.NT, ext, sv, z,
                        c->fsteps, (int)TP.fsteps.size(), fslots, c->bsteps,
                        (int)TP.bsteps.size(), bslots, c->st);
    } else {
      // sparse part forward, D u, sparse part backward
      const DenseArgs<T> da = dense_args<T>(c);
      const int NT1 = (int)TP.NT1;
      T* u = NT1 > 0 ? da.dvec : ext;
      if (NT1 > 0)
        LAPGPU_DISPATCH_E(TP.elog, hipLaunchKernelGGL, (k_tail_fwd<T, S, E_>), dim3(1),
                          dim3(kTailThreads), 0, c->stream, (int)TP.T0, (int)TP.NT, NT1, ext, sv, z,
                          da.dvec, c->fsteps, (int)TP.fsteps.size(), fslots, c->st);
      const int waves = (int)((DB.K + kDenseRows - 1) / kDenseRows);
      if (c->prof_sweeps) (void)hipEventRecord(c->ev[8], c->stream);
      hipLaunchKernelGGL(k_dense<T>, dim3((waves + kBlock / kWave - 1) / (kBlock / kWave)),
                         dim3(kBlock), 0, c->stream, (int)DB.K, (long)c->ldd, da.D, u, z + DB.T1,
                         c->st);
      if (c->prof_sweeps) (void)hipEventRecord(c->ev[9], c->stream);
      if (NT1 > 0)
        LAPGPU_DISPATCH_E(TP.elog, hipLaunchKernelGGL, (k_tail_bwd<T, S, E_>), dim3(1),
                          dim3(kTailThreads), 0, c->stream, (int)TP.T0, (int)TP.NT, NT1, z,
                          c->bsteps, (int)TP.bsteps.size(), bslots, c->st);
    }
  } else if (c->prof_sweeps) {
    (void)hipEventRecord(c->ev[6], c->stream);
  }
  if (c->prof_sweeps) (void)hipEventRecord(c->ev[7], c->stream);
  for (int k = (int)c->groups.size() - 1; k >= 0; --k) {
    const auto& g = c->groups[k];
    const int r = chain ? c->group_run[k] : -1;
    if (r >= 0) {
      if (k == c->runs[r].second - 1)
        launch_chain<T, T>(c, c->chainB, c->tkB[r], nruns + r, c->gBptr, c->gBidx, gBw, z, y, z);
      continue;
    }
    const int* hv = c->heavy + g.hoff_b;
    LAPGPU_DISPATCH_G(g.Gb, (launch_group<G_, T, T>), c, g.lo, g.hi, hv, g.nhb, c->gBptr,
                      c->gBidx, gBw, z, y, z);
  }
}

// Enqueued before the final synchronisation of a solve or apply: the
// timeout flag to the host.
int chain_fetch(lapgpu_ctx* c) {
  c->hchain[0] = 0;
  if (!c->chain_ctl) return 0;
  HIPC(c, hipMemcpyAsync(c->hchain, c->chain_ctl, sizeof(int32_t), hipMemcpyDeviceToHost,
                         c->stream));
  return 0;
}

// After that synchronisation: a wait of a chained run ran out (the call's
// result is not valid).  The context goes back to one launch per level
// group for good and the call reports LAPGPU_CHAIN_RETRY, which the entry
// points answer by running it again where they still hold its input.
constexpr int kChainRetry = LAPGPU_CHAIN_RETRY;
void drop_graphs(lapgpu_ctx* c);
int chain_check(lapgpu_ctx* c) {
  if (c->hchain[0] == 0) return 0;
  c->hchain[0] = 0;
  c->use_chain = false;
  drop_graphs(c);
  HIPC(c, hipMemsetAsync(c->chain_ctl, 0, sizeof(int), c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return set_err(c, kChainRetry,
                 "a chained run of level groups timed out waiting for a value; the context now "
                 "runs one launch per level group: solve again");
}

template <int G>
void launch_sweeps_mode(lapgpu_ctx* c, const double* rin) {
  if (c->precision == LAPGPU_FP32_INNER)
    launch_sweeps<float, TailSlotF>(c, rin, c->y32, c->z32, c->gFw32, c->gBw32, c->s32,
                                    c->Ew32, c->ext32, c->fslots32, c->bslots32);
  else
    launch_sweeps<double, TailSlotD>(c, rin, c->y, c->zF, c->gFw, c->gBw, c->s, c->Ew, c->ext,
                                     c->fslots, c->bslots);
}

// y (or y32) = forward, diagonal and backward sweeps of r (no mean removal)
void sweeps(lapgpu_ctx* c, const double* rin) {
  LAPGPU_DISPATCH_G(c->G_ldl, launch_sweeps_mode<G_>, c, rin);
}

// mean removal and z.r (sum|z|, sum|p| every 10th iteration) after sweeps
void finish_precond(lapgpu_ctx* c, const double* rin) {
  const int n = (int)c->n;
  if (c->precision == LAPGPU_FP32_INNER)
    hipLaunchKernelGGL(k_vsum<float>, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n,
                       c->z32, c->part, c->st);
  else
    hipLaunchKernelGGL(k_vsum<double>, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n,
                       c->zF, c->part, c->st);
  hipLaunchKernelGGL(k_scalar_mu, dim3(1), dim3(kScalarThreads), 0, c->stream, n, c->part,
                     c->st);
  if (c->precision == LAPGPU_FP32_INNER)
    hipLaunchKernelGGL(k_zdot<float>, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n,
                       c->z32, rin, c->p, c->z, c->part, c->st);
  else
    hipLaunchKernelGGL(k_zdot<double>, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n,
                       c->zF, rin, c->p, c->z, c->part, c->st);
}

// z = LDLsolver(r) = sweeps then mean removal; partial z.r etc.
void precond(lapgpu_ctx* c, const double* rin) {
  sweeps(c, rin);
  finish_precond(c, rin);
}

void drop_graphs(lapgpu_ctx* c) {
  if (c->iter_exec) (void)hipGraphExecDestroy(c->iter_exec);
  if (c->iter_exec_rep) (void)hipGraphExecDestroy(c->iter_exec_rep);
  c->iter_exec = c->iter_exec_rep = nullptr;
  for (auto& e : c->m.exec) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
}

// the multi-RHS workspace and tail CSRs (rebuilt on the next multi solve)
void free_multi(lapgpu_ctx* c) {
  auto& m = c->m;
  for (auto& e : m.exec)
    if (e) (void)hipGraphExecDestroy(e);
  void* dp[] = {m.r, m.x, m.bestx, m.p, m.q, m.y, m.z, m.io, m.part, m.ext,
                m.y32, m.z32, m.pf, m.ext32, m.r0, m.flags, m.st,
                m.dvec, m.dpart, m.dvec32, m.dpart32};
  for (void* a : dp) dev_free(c, a);
  // the host mirror may still be the target of a copy on the stream
  if (m.hst) (void)hipStreamSynchronize(c->stream);
  lapgpu::pinned_put(m.hst, sizeof(DevState) * kMultiMax);
  m = lapgpu_ctx::Multi();
}

void free_f32(lapgpu_ctx* c) {
  void* fp[] = {c->Lval32, c->gFw32, c->gBw32, c->s32, c->Ew32, c->y32, c->z32, c->ext32, c->pf,
                c->fslots32, c->bslots32, c->r0, c->D32, c->dvec32};
  for (void* a : fp) dev_free(c, a);
  c->Lval32 = c->gFw32 = c->gBw32 = c->s32 = c->Ew32 = c->y32 = c->z32 = c->ext32 = c->pf = nullptr;
  c->D32 = c->dvec32 = nullptr;
  c->fslots32 = c->bslots32 = nullptr;
  c->r0 = nullptr;
  c->f32_ready = false;
}

int free_graph(lapgpu_ctx* c) {
  drop_graphs(c);
  free_f32(c);
  free_multi(c);
  void* tp[] = {c->Eptr, c->Eidx, c->Ew, c->ext, c->fsteps, c->bsteps, c->fslots, c->bslots,
                c->D, c->dvec};
  for (void* a : tp) dev_free(c, a);
  c->Eptr = c->Eidx = nullptr;
  c->Ew = c->ext = nullptr;
  c->D = c->dvec = nullptr;
  c->ldd = 0;
  c->fsteps = c->bsteps = nullptr;
  c->fslots = c->bslots = nullptr;
  int* ip[] = {c->perm, c->Lptr, c->Lidx, c->lvl, c->gFptr, c->gFidx, c->gBptr, c->gBidx,
               c->heavy, c->mheavy};
  for (int* a : ip) dev_free(c, a);
  void* cp[] = {c->chainF, c->chainB, c->chain_ctl};
  for (void* a : cp) dev_free(c, a);
  c->chainF = c->chainB = nullptr;
  c->chain_ctl = nullptr;
  c->runs.clear();
  c->group_run.clear();
  double* dp[] = {c->Lval, c->s, c->b, c->r, c->x, c->bestx, c->p, c->q, c->y, c->z, c->io,
                  c->gFw, c->gBw, c->zF};
  for (double* a : dp) dev_free(c, a);
  c->perm = c->Lptr = c->Lidx = c->lvl = nullptr;
  c->gFptr = c->gFidx = c->gBptr = c->gBidx = c->heavy = c->mheavy = nullptr;
  c->Lval = c->s = c->b = c->r = c->x = c->bestx = c->p = c->q = c->y = c->z = c->io = nullptr;
  c->gFw = c->gBw = c->zF = nullptr;
  c->groups.clear();
  return 0;
}

template <typename D, typename S>
std::vector<D> narrow(const std::vector<S>& v) {
  std::vector<D> o(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = (D)v[i];
  return o;
}

// FP32 copies of la, the sweep weights and the tail slots; FP32 sweep and p
// buffers; r0 for the FP64 residual replacement.
int prepare_f32(lapgpu_ctx* c) {
  if (c->f32_ready) return 0;
  const Plan& P = c->plan;
  const int64_t n = P.n;
  int rc;
  if ((rc = dev_upload(c, &c->Lval32, narrow<float>(P.Lval)))) return rc;
  if ((rc = dev_upload(c, &c->gFw32, narrow<float>(P.blocks.fw)))) return rc;
  if ((rc = dev_upload(c, &c->gBw32, narrow<float>(P.blocks.bw)))) return rc;
  if ((rc = dev_alloc(c, &c->z32, (size_t)n))) return rc;
  if ((rc = dev_upload(c, &c->s32, narrow<float>(P.s)))) return rc;
  if ((rc = dev_upload(c, &c->Ew32, narrow<float>(P.tail.Ew)))) return rc;
  auto slots32 = [](const std::vector<TailSlot>& v) {
    std::vector<TailSlotF> o(v.size());
    for (size_t i = 0; i < v.size(); ++i) o[i] = TailSlotF{(float)v[i].w, v[i].src, v[i].meta};
    return o;
  };
  if ((rc = dev_upload(c, &c->fslots32, slots32(P.tail.fslots)))) return rc;
  if ((rc = dev_upload(c, &c->bslots32, slots32(P.tail.bslots)))) return rc;
  if ((rc = dev_alloc(c, &c->y32, (size_t)n))) return rc;
  if ((rc = dev_alloc(c, &c->pf, (size_t)n))) return rc;
  if ((rc = dev_alloc(c, &c->r0, (size_t)n))) return rc;
  if ((rc = dev_zeros(c, &c->ext32, ext_len(c)))) return rc;
  if (P.dense.K > 0) {
    const size_t cnt = (size_t)P.dense.K * (size_t)c->ldd;
    if ((rc = dev_alloc(c, &c->D32, cnt))) return rc;
    hipLaunchKernelGGL(k_narrow, dim3(2048), dim3(kBlock), 0, c->stream, cnt, c->D, c->D32);
    HIPC(c, hipGetLastError());
    if ((rc = dev_zeros(c, &c->dvec32, (size_t)c->ldd))) return rc;
  }
  c->f32_ready = true;
  return 0;
}

// D of the dense block: its programs are uploaded for the build only and
// run on every unit vector, one workgroup per column (columns in chunks of
// kBasisChunk per launch); the FP32 copy follows in prepare_f32.
constexpr int kBasisChunk = 2048;

int build_dense(lapgpu_ctx* c) {
  const DenseBlock& DB = c->plan.dense;
  if (DB.K == 0) return 0;
  int rc;
  const size_t cnt = (size_t)DB.K * (size_t)c->ldd;
  if ((rc = dev_alloc(c, &c->D, cnt))) return rc;
  if ((rc = dev_zeros(c, &c->dvec, (size_t)c->ldd))) return rc;
  std::vector<TailStepD> fs(DB.fsteps.size()), bs(DB.bsteps.size());
  std::memcpy(fs.data(), DB.fsteps.data(), sizeof(TailStep) * fs.size());
  std::memcpy(bs.data(), DB.bsteps.data(), sizeof(TailStep) * bs.size());
  std::vector<TailSlotD> fl(DB.fslots.size()), bl(DB.bslots.size());
  std::memcpy(fl.data(), DB.fslots.data(), sizeof(TailSlot) * fl.size());
  std::memcpy(bl.data(), DB.bslots.data(), sizeof(TailSlot) * bl.size());
  TailStepD *dfs = nullptr, *dbs = nullptr;
  TailSlotD *dfl = nullptr, *dbl = nullptr;
  if ((rc = dev_upload(c, &dfs, fs)) || (rc = dev_upload(c, &dbs, bs)) ||
      (rc = dev_upload(c, &dfl, fl)) || (rc = dev_upload(c, &dbl, bl)))
    return rc;
  // P.s in position order: the block's scales are s[T1 ...] (uploaded first)
  for (int64_t j0 = 0; j0 < DB.K; j0 += kBasisChunk) {
    const int nb = (int)std::min<int64_t>(kBasisChunk, DB.K - j0);
    LAPGPU_DISPATCH_E(DB.elog, hipLaunchKernelGGL, (k_dense_basis<E_>), dim3(nb), dim3(kTailThreads),
                      0, c->stream, (int)DB.K, (int)DB.T1, (int)j0, c->s, dfs, (int)fs.size(), dfl,
                      dbs, (int)bs.size(), dbl, c->D, (long)c->ldd);
    HIPC(c, hipGetLastError());
  }
  dev_free(c, dfs);
  dev_free(c, dbs);
  dev_free(c, dfl);
  dev_free(c, dbl);
  HIPC(c, hipStreamSynchronize(c->stream));
  return 0;
}

// Runs of narrow groups and their single-RHS tickets: per group of a run,
// in sweep order, its heavy rows four at a time (a wave each), then its
// light rows in ranges of whole passes of the workgroup (as many passes as
// keep about kChainTickets tickets per group, at most kChainPasses).

int build_chain(lapgpu_ctx* c) {
  const int ng = (int)c->groups.size();
  c->group_run.assign(ng, -1);
  c->runs.clear();
  c->tkF.clear();
  c->tkB.clear();
  for (int g = 0; g < ng;) {
    int e = g;
    while (e < ng && c->groups[e].hi - c->groups[e].lo <= c->chain_narrow) ++e;
    if (e - g >= std::max(2, c->chain_min_run)) {
      for (int k = g; k < e; ++k) c->group_run[k] = (int)c->runs.size();
      c->runs.push_back({g, e});
    }
    g = std::max(e, g + 1);
  }
  c->chain_lo = c->chain_hi = 0;
  if (c->runs.empty()) return 0;
  c->chain_lo = c->groups[c->runs.front().first].lo;
  c->chain_hi = c->groups[c->runs.back().second - 1].hi;
  std::vector<ChainTicket> tf, tb;
  auto add = [&](std::vector<ChainTicket>& out, const lapgpu_ctx::GroupLaunch& g, int G, int nh,
                 int hoff) {
    constexpr int hpt = kBlock / kWave;
    for (int h = 0; h < nh; h += hpt) out.push_back({hoff + h, hoff + std::min(nh, h + hpt), 0, 0});
    const int rpp = kBlock / G;
    const int64_t passes = ((int64_t)(g.hi - g.lo) + rpp - 1) / rpp;
    const int per = (int)std::min<int64_t>(kChainPasses, std::max<int64_t>(1, passes / kChainTickets));
    const int span = per * rpp;
    for (int lo = g.lo; lo < g.hi; lo += span) out.push_back({lo, std::min(g.hi, lo + span), G, 0});
  };
  for (const auto& r : c->runs) {
    const int f0 = (int)tf.size();
    for (int k = r.first; k < r.second; ++k) {
      const auto& g = c->groups[k];
      add(tf, g, g.Gf, g.nhf, g.hoff_f);
    }
    c->tkF.push_back({f0, (int)tf.size()});
    const int b0 = (int)tb.size();
    for (int k = r.second - 1; k >= r.first; --k) {
      const auto& g = c->groups[k];
      add(tb, g, g.Gb, g.nhb, g.hoff_b);
    }
    c->tkB.push_back({b0, (int)tb.size()});
  }
  int rc;
  if ((rc = dev_upload(c, &c->chainF, tf))) return rc;
  if ((rc = dev_upload(c, &c->chainB, tb))) return rc;
  std::vector<int> ctl(2 + 2 * c->runs.size(), 0);
  ctl[1] = c->chain_poll_max;
  return dev_upload(c, &c->chain_ctl, ctl);
}

// The level groups before the tail: composite rows, the per-group lane
// count and the heavy-row lists of both sweeps.
int upload_groups(lapgpu_ctx* c, bool blocks_ready) {
  Plan& P = c->plan;
  if (!blocks_ready) build_blocks(P, c->block_depth, c->block_growth, c->block_extra);
  const LevelBlocks& B = P.blocks;
  c->groups.clear();
  std::vector<int32_t> heavy, mheavy;
  // multi-RHS light rows: slices so that a lane sums at most ~8 entries
  auto plan_multi = [&](const std::vector<int32_t>& ptr, int lo, int hi, int& S, int& nh,
                        int& hoff) {
    int64_t rows = 0, ent = 0;
    hoff = (int)mheavy.size();
    for (int v = lo; v < hi; ++v) {
      const int64_t len = ptr[v + 1] - ptr[v];
      if (len > kMHeavyLen) {
        mheavy.push_back(v);
      } else {
        ++rows;
        ent += len;
      }
    }
    nh = (int)mheavy.size() - hoff;
    const double avg = rows ? (double)ent / (double)rows : 1.0;
    S = 1;
    while (S < 64 && (double)S * 8.0 < avg) S *= 2;
  };
  auto plan_sweep = [&](const std::vector<int32_t>& ptr, int lo, int hi, int& G, int& nh,
                        int& hoff) {
    int64_t light_rows = 0, light_ent = 0;
    int64_t maxdeg = 0;
    for (int v = lo; v < hi; ++v) maxdeg = std::max<int64_t>(maxdeg, ptr[v + 1] - ptr[v]);
    // lanes per row from the mean of the rows that are not heavy
    int64_t tot = ptr[hi] - ptr[lo];
    G = pick_group(hi > lo ? (double)tot / (double)(hi - lo) : 1.0);
    if (G > 32) G = 32;
    const int64_t lim = 8 * (int64_t)G;
    hoff = (int)heavy.size();
    for (int v = lo; v < hi; ++v)
      if (ptr[v + 1] - ptr[v] > lim) heavy.push_back(v);
      else {
        ++light_rows;
        light_ent += ptr[v + 1] - ptr[v];
      }
    nh = (int)heavy.size() - hoff;
    (void)maxdeg;
  };
  for (size_t g = 0; g + 1 < B.lev.size(); ++g) {
    lapgpu_ctx::GroupLaunch gl{};
    gl.lo = P.lvl[B.lev[g]];
    gl.hi = P.lvl[B.lev[g + 1]];
    if (gl.hi <= gl.lo) continue;
    plan_sweep(B.fptr, gl.lo, gl.hi, gl.Gf, gl.nhf, gl.hoff_f);
    plan_sweep(B.bptr, gl.lo, gl.hi, gl.Gb, gl.nhb, gl.hoff_b);
    plan_multi(B.fptr, gl.lo, gl.hi, gl.Sf, gl.mnhf, gl.mhoff_f);
    plan_multi(B.bptr, gl.lo, gl.hi, gl.Sb, gl.mnhb, gl.mhoff_b);
    c->groups.push_back(gl);
  }
  int rc;
  if ((rc = dev_upload(c, &c->gFptr, B.fptr))) return rc;
  if ((rc = dev_upload(c, &c->gFidx, B.fidx))) return rc;
  if ((rc = dev_upload(c, &c->gFw, B.fw))) return rc;
  if ((rc = dev_upload(c, &c->gBptr, B.bptr))) return rc;
  if ((rc = dev_upload(c, &c->gBidx, B.bidx))) return rc;
  if ((rc = dev_upload(c, &c->gBw, B.bw))) return rc;
  if ((rc = dev_upload(c, &c->heavy, heavy))) return rc;
  if ((rc = dev_upload(c, &c->mheavy, mheavy))) return rc;
  return build_chain(c);
}

// blocks_ready: P.blocks already holds the level groups (a loaded plan)
int upload_plan(lapgpu_ctx* c, bool blocks_ready = false) {
  Plan& P = c->plan;
  const int64_t n = P.n;
  int rc;
  // the multi-RHS workspace is sized by the previous plan (n * K vectors,
  // NT * K tail inputs): drop it, the next multi-RHS solve sizes it again
  free_multi(c);
  if ((rc = dev_upload(c, &c->perm, P.perm))) return rc;
  if ((rc = dev_upload(c, &c->Lptr, P.Lptr))) return rc;
  if ((rc = dev_upload(c, &c->Lidx, P.Lidx))) return rc;
  if ((rc = dev_upload(c, &c->Lval, P.Lval))) return rc;
  if ((rc = dev_upload(c, &c->s, P.s))) return rc;
  if ((rc = upload_groups(c, blocks_ready))) return rc;
  if ((rc = dev_upload(c, &c->lvl, P.lvl))) return rc;
  if ((rc = dev_upload(c, &c->Eptr, P.tail.Eptr))) return rc;
  if ((rc = dev_upload(c, &c->Eidx, P.tail.Eidx))) return rc;
  if ((rc = dev_upload(c, &c->Ew, P.tail.Ew))) return rc;
  c->ldd = P.dense.K > 0 ? (P.dense.K + 255) / 256 * 256 : 0;
  if ((rc = dev_zeros(c, &c->ext, ext_len(c)))) return rc;
  if ((rc = build_dense(c))) return rc;
  {
    static_assert(sizeof(TailStepD) == sizeof(TailStep), "step layout");
    static_assert(sizeof(TailSlotD) == sizeof(TailSlot), "slot layout");
    std::vector<TailStepD> fs(P.tail.fsteps.size()), bs(P.tail.bsteps.size());
    std::memcpy(fs.data(), P.tail.fsteps.data(), sizeof(TailStep) * fs.size());
    std::memcpy(bs.data(), P.tail.bsteps.data(), sizeof(TailStep) * bs.size());
    std::vector<TailSlotD> fl(P.tail.fslots.size()), bl(P.tail.bslots.size());
    std::memcpy(fl.data(), P.tail.fslots.data(), sizeof(TailSlot) * fl.size());
    std::memcpy(bl.data(), P.tail.bslots.data(), sizeof(TailSlot) * bl.size());
    if ((rc = dev_upload(c, &c->fsteps, fs))) return rc;
    if ((rc = dev_upload(c, &c->bsteps, bs))) return rc;
    if ((rc = dev_upload(c, &c->fslots, fl))) return rc;
    if ((rc = dev_upload(c, &c->bslots, bl))) return rc;
  }
  drop_graphs(c);
  free_f32(c);
  if (c->precision == LAPGPU_FP32_INNER) {
    int rc2 = prepare_f32(c);
    if (rc2) return rc2;
  }
  double** ws[] = {&c->b, &c->r, &c->x, &c->bestx, &c->p, &c->q, &c->y, &c->z, &c->io, &c->zF};
  for (double** w : ws)
    if ((rc = dev_alloc(c, w, (size_t)n))) return rc;
  c->n = n;
  const double avgL = (double)P.Lidx.size() / (double)n;
  const double avgF = (double)P.nnzF / (double)n;
  c->G_spmv = pick_group(avgL);
  c->G_ldl = pick_group(avgF);
  if (const char* e = getenv("LAPGPU_G_SPMV")) c->G_spmv = atoi(e);
  if (const char* e = getenv("LAPGPU_G_LDL")) c->G_ldl = atoi(e);
  return 0;
}

// Identity plan: la only, no factor (lapgpu_spmv without an LDLinv).
void identity_plan(lapgpu_ctx* c) {
  Plan& P = c->plan;
  const int64_t n = c->L.n;
  P = Plan();
  P.n = n;
  P.nnzF = 0;
  P.perm.resize(n);
  P.iperm.resize(n);
  for (int64_t i = 0; i < n; ++i) P.perm[i] = P.iperm[i] = (int32_t)i;
  P.Lptr = c->L.ptr;
  P.Lidx = c->L.idx;
  P.Lval = c->L.val;
  P.Fptr.assign(n + 1, 0);
  P.Bptr.assign(n + 1, 0);
  P.s.assign(n, 1.0);
  P.nlev = 1;
  P.lvl = {0, (int32_t)n};
  P.tail_level = 1;
  P.tail = TailProgram();
}

void fill_state(const lapgpu_ctx* c, DevState& h, double tol, int64_t maxits,
                double maxtime, int stag_test, int flags, bool fp32) {
  std::memset(&h, 0, sizeof(DevState));
  h.tol = tol;
  h.maxits = maxits;
  h.stag_test = stag_test;
  h.stag_fac = 1.0 - 1.0 / (double)stag_test;  // -Inf for stag_test == 0
  h.center = (flags & LAPGPU_NO_CENTER) ? 0 : 1;
  h.fp32 = fp32 ? 1 : 0;
  h.switch_tol = c->switch_tol;
  h.has_maxtime = (maxtime > 0 && std::isfinite(maxtime)) ? 1 : 0;
  h.maxticks = h.has_maxtime ? (uint64_t)(maxtime * c->wallclock_khz * 1e3) : 0;
  h.bestnr = 1.0;
  h.nr = 1.0;
}

void init_state(lapgpu_ctx* c, double tol, int64_t maxits, double maxtime,
                int stag_test, int flags) {
  fill_state(c, *c->hst, tol, maxits, maxtime, stag_test, flags,
             c->precision == LAPGPU_FP32_INNER);
}

const char* stop_message(int reason) {
  switch (reason) {
    case 0: return "PCG stopped due to small nr (norm(r)/nb)";
    case 2: return "PCG New stopped at maxtime.";
    case 3: return "PCG Stopped by stagnation test";
    case 4: return "PCG: Stopped due to stagnation: al*sum(abs.(p)) < eps(Tval)*sum(abs.(x))";
    case 5: return "PCG: Stopped due to stagnation: al*sum(abs.(q)) < eps(Tval)*sum(abs.(r))";
    case 6: return "PCG: Stopped due to stagnation: beta*sum(abs.(p)) < eps(Tval)*sum(abs.(z))";
    case 7: return "PCG Stopped due to large rho";
    case 8: return "PCG Stopped due to large beta";
    default: return "";
  }
}

// one PCG iteration (src/pcg.jl:288-431); `replace`: FP32-mode iteration
// that ends with an FP64 residual replacement
void enqueue_iteration(lapgpu_ctx* c, bool prof, bool replace) {
  const int n = (int)c->n;
  const bool f32 = c->precision == LAPGPU_FP32_INNER;
  if (prof) (void)hipEventRecord(c->ev[0], c->stream);
  spmv(c, c->p, c->q);
  if (prof) (void)hipEventRecord(c->ev[1], c->stream);
  hipLaunchKernelGGL(k_scalar_al, dim3(1), dim3(kScalarThreads), 0, c->stream, c->part, c->st);
  hipLaunchKernelGGL(k_update, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n,
                     c->x, c->bestx, c->p, c->q, c->r, c->part, c->st);
  hipLaunchKernelGGL(k_scalar_nr, dim3(1), dim3(kScalarThreads), 0, c->stream, c->part, c->st);
  if (f32) {
    LAPGPU_DISPATCH_G(c->G_spmv, launch_resid<G_>, c, replace ? 1 : 0);
    hipLaunchKernelGGL(k_scalar_resid, dim3(1), dim3(kScalarThreads), 0, c->stream,
                       replace ? 1 : 0, c->part, c->st);
  }
  if (prof) (void)hipEventRecord(c->ev[2], c->stream);
  c->prof_sweeps = prof;
  sweeps(c, c->r);
  c->prof_sweeps = false;
  if (prof) (void)hipEventRecord(c->ev[3], c->stream);
  finish_precond(c, c->r);
  hipLaunchKernelGGL(k_scalar_rho, dim3(1), dim3(kScalarThreads), 0, c->stream, c->part, c->st);
  if (f32)
    hipLaunchKernelGGL(k_pupdate_f, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n,
                       c->z, c->p, c->pf, c->st);
  else
    hipLaunchKernelGGL(k_pupdate, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n,
                       c->z, c->p, c->st);
  if (prof) (void)hipEventRecord(c->ev[4], c->stream);
}

// The running-solve window of lapgpu_cancel: k columns, flags cleared.
struct RunGuard {
  lapgpu_ctx* c;
  RunGuard(lapgpu_ctx* c_, int64_t k) : c(c_) {
    std::lock_guard<std::mutex> g(c->cancel_mu);
    c->cancel_cols.assign((size_t)k, 0);
    c->cancel_any.store(false);
    c->running = true;
  }
  ~RunGuard() {
    std::lock_guard<std::mutex> g(c->cancel_mu);
    c->running = false;
    c->cancel_any.store(false);
    c->cancel_cols.clear();
  }
};

// Columns [base, base + k) cancelled since the last call, as a bit mask
// (cleared as they are taken; flags of earlier columns, already solved,
// are dropped).
uint64_t take_cancel(lapgpu_ctx* c, int64_t base, int k) {
  if (!c->cancel_any.load(std::memory_order_acquire)) return 0;
  std::lock_guard<std::mutex> g(c->cancel_mu);
  uint64_t mask = 0;
  bool left = false;
  for (size_t j = 0; j < c->cancel_cols.size(); ++j) {
    if (!c->cancel_cols[j]) continue;
    const int64_t r = (int64_t)j - base;
    if (r < k) {
      if (r >= 0) mask |= 1ull << r;
      c->cancel_cols[j] = 0;
    } else {
      left = true;
    }
  }
  c->cancel_any.store(left);
  return mask;
}

int solve_device(lapgpu_ctx* c, const double* d_b, double* d_x, double tol,
                 int64_t maxits, double maxtime, int stag_test, int flags,
                 int64_t* its_out, double* relres_out, int* stop_out) {
  if (!c->has_F) return set_err(c, -2, "solve: upload the Laplacian and LDLinv first");
  const auto t_start = std::chrono::steady_clock::now();
  const bool prof = (flags & LAPGPU_PROFILE) != 0;
  const bool verbose = (flags & LAPGPU_VERBOSE) != 0;
  const int n = (int)c->n;
  init_state(c, tol, maxits, maxtime, stag_test, flags);
  HIPC(c, hipMemcpyAsync(c->st, c->hst, sizeof(DevState), hipMemcpyHostToDevice, c->stream));

  // ---- setup: r = b - mean(b), z = pre(r), p = z, rho = r.z ---------------
  hipLaunchKernelGGL(k_gather_perm, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n,
                     c->perm, d_b, c->b, c->part);
  hipLaunchKernelGGL(k_scalar_mean, dim3(1), dim3(kScalarThreads), 0, c->stream, n, c->part, c->st);
  hipLaunchKernelGGL(k_center, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n, c->b,
                     c->r, c->x, c->bestx, c->part, c->st);
  hipLaunchKernelGGL(k_scalar_nb, dim3(1), dim3(kScalarThreads), 0, c->stream, c->part, c->st);
  const bool f32 = c->precision == LAPGPU_FP32_INNER;
  if (f32)
    HIPC(c, hipMemcpyAsync(c->r0, c->r, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  precond(c, c->r);
  hipLaunchKernelGGL(k_scalar_rho0, dim3(1), dim3(kScalarThreads), 0, c->stream, c->part, c->st);
  if (f32)
    hipLaunchKernelGGL(k_copy_f, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n, c->z,
                       c->p, c->pf, c->st);
  else
    hipLaunchKernelGGL(k_copy, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n, c->z,
                       c->p, c->st);
  HIPC(c, hipGetLastError());

  // the setup may already have stopped (b == 0, maxits == 0)
  HIPC(c, hipMemcpyAsync(c->hst, c->st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, sync_stream(c));
  if (verbose) {
    printf("<b, LDLi b> is %.17g\n", c->hst->nb == 0.0 ? 0.0 : c->hst->rho0);
    fflush(stdout);
  }

  // ---- iterations: poll the stop flag one iteration behind ---------------
  double acc_spmv = 0, acc_ldl = 0, acc_vec = 0, acc_fg = 0, acc_ext = 0, acc_tail = 0,
         acc_bg = 0, acc_tf = 0, acc_dn = 0, acc_tb = 0;
  int64_t enq = 0, polled = 0;
  bool done = c->hst->stopped != 0;
  const int rep_every = f32 ? c->replace_every : 0;
  const bool captured = !prof && c->use_graph && !(flags & LAPGPU_EAGER);
  if (captured) {
    for (int v = 0; v < (f32 ? 2 : 1); ++v) {
      hipGraphExec_t* ex = v ? &c->iter_exec_rep : &c->iter_exec;
      if (*ex) continue;
      LAPGPU_LOCK();
      hipGraph_t g = nullptr;
      HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      enqueue_iteration(c, false, v == 1);
      HIPC(c, hipStreamEndCapture(c->stream, &g));
      hipError_t ie = hipGraphInstantiate(ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPC(c, ie);
    }
  }
  while (!done) {
    // iteration number enq + 1 of this solve (itcnt on the device)
    const bool replace = rep_every > 0 && ((enq + 1) % rep_every == 0);
    hipGraphExec_t ex = replace ? c->iter_exec_rep : c->iter_exec;
    if (captured && ex)
      HIPC(c, hipGraphLaunch(ex, c->stream));
    else
      enqueue_iteration(c, prof, replace);
    const int slot = (int)(enq % lapgpu_ctx::kRing);
    HIPC(c, hipMemcpyAsync(&c->hflags[slot], &c->st->stopped, sizeof(int32_t),
                           hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipEventRecord(c->ring_ev[slot], c->stream));
    ++enq;
    if (prof) {
      HIPC(c, hipEventSynchronize(c->ev[4]));
      float a = 0, b2 = 0, d = 0, e = 0;
      (void)hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
      (void)hipEventElapsedTime(&b2, c->ev[2], c->ev[3]);
      (void)hipEventElapsedTime(&d, c->ev[1], c->ev[2]);
      (void)hipEventElapsedTime(&e, c->ev[3], c->ev[4]);
      acc_spmv += a;
      acc_ldl += b2;
      acc_vec += d + e;
      float f1 = 0, f2 = 0, f3 = 0, f4 = 0;
      (void)hipEventElapsedTime(&f1, c->ev[2], c->ev[5]);
      (void)hipEventElapsedTime(&f2, c->ev[5], c->ev[6]);
      (void)hipEventElapsedTime(&f3, c->ev[6], c->ev[7]);
      (void)hipEventElapsedTime(&f4, c->ev[7], c->ev[3]);
      acc_fg += f1;
      acc_ext += f2;
      acc_tail += f3;
      acc_bg += f4;
      if (c->plan.dense.K > 0) {
        float g1 = 0, g2 = 0, g3 = 0;
        (void)hipEventElapsedTime(&g1, c->ev[6], c->ev[8]);
        (void)hipEventElapsedTime(&g2, c->ev[8], c->ev[9]);
        (void)hipEventElapsedTime(&g3, c->ev[9], c->ev[7]);
        acc_tf += g1;
        acc_dn += g2;
        acc_tb += g3;
      }
    }
    // keep one iteration in flight: wait for the flags of all older ones
    while (!done && polled < enq - 1) {
      const int ps = (int)(polled % lapgpu_ctx::kRing);
      HIPC(c, wait_event(c, c->ring_ev[ps]));
      ++polled;
      if (c->hflags[ps]) done = true;
    }
    // cancelled: stop after the iterations already enqueued
    if (!done && take_cancel(c, 0, 1)) {
      hipLaunchKernelGGL(k_cancel, dim3(1), dim3(kWave), 0, c->stream, c->st, 1, 1ull);
      done = true;
    }
  }
  hipLaunchKernelGGL(k_final_copy, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n,
                     c->x, c->bestx, c->st);
  LAPGPU_DISPATCH_G(c->G_spmv, launch_true_resid<G_>, c);
  hipLaunchKernelGGL(k_scatter_perm, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, n,
                     c->perm, c->bestx, d_x);
  HIPC(c, hipMemcpyAsync(c->hst, c->st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
  if (int rc = chain_fetch(c)) return rc;
  HIPC(c, sync_stream(c));
  HIPC(c, hipGetLastError());
  if (int rc = chain_check(c)) return rc;
  const DevState& h = *c->hst;
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  if (its_out) *its_out = h.itcnt;
  if (relres_out) *relres_out = h.nr;
  if (stop_out) *stop_out = h.stop_reason;
  c->last_true.assign(1, h.true_nr);
  if (prof && h.itcnt > 0) {
    c->t_spmv = acc_spmv / (double)enq;
    c->t_ldl = acc_ldl / (double)enq;
    c->t_vec = acc_vec / (double)enq;
    c->t_fgroups = acc_fg / (double)enq;
    c->t_tail_ext = acc_ext / (double)enq;
    c->t_tail = acc_tail / (double)enq;
    c->t_bgroups = acc_bg / (double)enq;
    c->t_tail_fwd = acc_tf / (double)enq;
    c->t_dense = acc_dn / (double)enq;
    c->t_tail_bwd = acc_tb / (double)enq;
  }
  c->t_total = secs * 1e3;
  c->t_its = h.itcnt;
  if (h.stop_reason == 3) {
    // printed unconditionally by the reference (src/pcg.jl:378)
    printf("PCG Stopped by stagnation test %d\n", stag_test);
  } else if (verbose && h.stop_reason != 1 && h.stop_reason != 9) {
    if (h.stop_reason == 0)
      printf("%s = %.17g\nnb = %.17g\n", stop_message(0), h.nr, h.nb);
    else
      printf("%s\n", stop_message(h.stop_reason));
  }
  if (verbose && h.stop_reason != 9) {
    printf("PCG stopped after: %.3f seconds and %lld iterations with relative error %.17g.\n",
           secs, (long long)h.itcnt, h.nr);
    printf("rho = %.17g\n", h.rho);
  }
  fflush(stdout);
  c->err.clear();
  return 0;
}

// ---- k right-hand sides at once (lapgpu_multi.hpp) ---------------------------

// Calls F with the compile-time rhs count K_ = k (8, 16, 32 or 64).
#define LAPGPU_DISPATCH_K(k, F, ...)                                \
  switch (k) {                                                      \
    case 8: { constexpr int K_ = 8; F(__VA_ARGS__); break; }        \
    case 16: { constexpr int K_ = 16; F(__VA_ARGS__); break; }      \
    case 32: { constexpr int K_ = 32; F(__VA_ARGS__); break; }      \
    default: { constexpr int K_ = 64; F(__VA_ARGS__); break; }      \
  }

// the dense block for K columns in precision T: D, its inputs and the
// split partials
template <typename T>
struct MDenseArgs {
  const T* D;
  T* dvec;
  T* part;
};
template <typename T>
MDenseArgs<T> mdense_args(lapgpu_ctx* c);
template <>
MDenseArgs<double> mdense_args<double>(lapgpu_ctx* c) { return {c->D, c->m.dvec, c->m.dpart}; }
template <>
MDenseArgs<float> mdense_args<float>(lapgpu_ctx* c) { return {c->D32, c->m.dvec32, c->m.dpart32}; }

// k splits of the multi-RHS dense product: enough workgroups for the chip
// (about four per CU) with at least two slabs per split
int dense_splits(const lapgpu_ctx* c) {
  const int64_t tiles = (c->plan.dense.K + kMDenseRows - 1) / kMDenseRows;
  int64_t ns = std::max<int64_t>(1, (1024 + tiles - 1) / tiles);
  ns = std::min<int64_t>(ns, std::max<int64_t>(1, c->ldd / (2 * kMDenseSlab)));
  return (int)std::min<int64_t>(ns, 16);
}

// Workspace for n * K elements per vector (kept while it is large enough);
// FP32 buffers only in the FP32-inner mode.
int multi_prepare(lapgpu_ctx* c, int K) {
  auto& m = c->m;
  const int64_t n = c->n;
  int rc;
  if (!m.part) {
    if ((rc = dev_alloc(c, &m.part, (size_t)kNumSlots * kMRed * kMultiMax))) return rc;
    if ((rc = dev_alloc(c, &m.st, (size_t)kMultiMax))) return rc;
    if ((rc = dev_alloc(c, &m.flags, 4))) return rc;
    HIPC(c, lapgpu::pinned_get((void**)&m.hst, sizeof(DevState) * kMultiMax));
  }
  const int64_t need = n * K;
  if (need > m.cap) {
    drop_graphs(c);
    double** ws[] = {&m.r, &m.x, &m.bestx, &m.p, &m.q, &m.y, &m.z, &m.io};
    for (double** w : ws)
      if ((rc = dev_alloc(c, w, (size_t)need))) return rc;
    if ((rc = dev_zeros(c, &m.ext, ext_len(c) * K))) return rc;
    if (c->plan.dense.K > 0) {
      if ((rc = dev_zeros(c, &m.dvec, (size_t)c->ldd * K))) return rc;
      if ((rc = dev_alloc(c, &m.dpart, (size_t)dense_splits(c) * c->plan.dense.K * K))) return rc;
    }
    m.cap = need;
  }
  if (c->precision == LAPGPU_FP32_INNER && need > m.cap32) {
    drop_graphs(c);
    float** ws[] = {&m.y32, &m.z32, &m.pf};
    for (float** w : ws)
      if ((rc = dev_alloc(c, w, (size_t)need))) return rc;
    if ((rc = dev_zeros(c, &m.ext32, ext_len(c) * K))) return rc;
    if ((rc = dev_alloc(c, &m.r0, (size_t)need))) return rc;
    if (c->plan.dense.K > 0) {
      if ((rc = dev_zeros(c, &m.dvec32, (size_t)c->ldd * K))) return rc;
      if ((rc = dev_alloc(c, &m.dpart32, (size_t)dense_splits(c) * c->plan.dense.K * K))) return rc;
    }
    m.cap32 = need;
  }
  return 0;
}

// One level group for K right-hand sides: light rows with S entry slices
// per row (S bounded by 64/K), rows longer than kMHeavyLen one workgroup
// each — in the same launch (its first nh workgroups) when K <= mgroup_merge_k,
// else in a launch of their own after the light rows.
template <int K, int S, int NT, typename T, typename TB>
void launch_mgroup_s(lapgpu_ctx* c, int lo, int hi, const int* heavy, int nh, const int* ptr,
                     const int* idx, const T* w, const T* A, const TB* B, T* out) {
  const int64_t threads = (int64_t)(hi - lo) * K * S;
  const int lead = NT == kMGroupThreads ? nh : 0;
  hipLaunchKernelGGL((k_mgroup<K, S, NT, T, TB>), dim3(lead + grid_for(threads, NT)), dim3(NT),
                     0, c->stream, (int)c->n, lo, hi, heavy, lead, ptr, idx, w, A, B, out,
                     c->m.flags);
  if (NT != kMGroupThreads && nh > 0)
    hipLaunchKernelGGL((k_mgroup_heavy<K, T, TB>), dim3(nh), dim3(kMGroupThreads), 0, c->stream,
                       (int)c->n, heavy, ptr, idx, w, A, B, out, c->m.flags);
}

template <int K, typename T, typename TB>
void launch_mgroup(lapgpu_ctx* c, int lo, int hi, int S, const int* heavy, int nh,
                   const int* ptr, const int* idx, const T* w, const T* A, const TB* B,
                   T* out) {
  const int s = std::min(S, kWave / K);
  if (K <= c->mgroup_merge_k) {
    constexpr int NT = kMGroupThreads;
    switch (s) {
      case 1: launch_mgroup_s<K, 1, NT>(c, lo, hi, heavy, nh, ptr, idx, w, A, B, out); break;
      case 2: launch_mgroup_s<K, (K <= 32 ? 2 : 1), NT>(c, lo, hi, heavy, nh, ptr, idx, w, A, B, out); break;
      case 4: launch_mgroup_s<K, (K <= 16 ? 4 : 1), NT>(c, lo, hi, heavy, nh, ptr, idx, w, A, B, out); break;
      default: launch_mgroup_s<K, (K <= 8 ? 8 : 1), NT>(c, lo, hi, heavy, nh, ptr, idx, w, A, B, out); break;
    }
  } else {
    constexpr int NT = kBlock;
    switch (s) {
      case 1: launch_mgroup_s<K, 1, NT>(c, lo, hi, heavy, nh, ptr, idx, w, A, B, out); break;
      case 2: launch_mgroup_s<K, (K <= 32 ? 2 : 1), NT>(c, lo, hi, heavy, nh, ptr, idx, w, A, B, out); break;
      case 4: launch_mgroup_s<K, (K <= 16 ? 4 : 1), NT>(c, lo, hi, heavy, nh, ptr, idx, w, A, B, out); break;
      default: launch_mgroup_s<K, (K <= 8 ? 8 : 1), NT>(c, lo, hi, heavy, nh, ptr, idx, w, A, B, out); break;
    }
  }
}

// z = LDLsolver(rin) without the mean removal for K interleaved right-hand
// sides: forward groups, the tail (one workgroup per column), backward
// groups.  T, S: sweep precision and tail slot type.
template <int K, typename T, typename S>
void msweeps_t(lapgpu_ctx* c, const double* rin, T* y, T* z, T* ext, const T* gFw, const T* gBw,
               const T* sv, const T* Ew, const S* fslots, const S* bslots) {
  auto& m = c->m;
  const Plan& P = c->plan;
  const int n = (int)c->n;
  for (const auto& g : c->groups)
    launch_mgroup<K, T, double>(c, g.lo, g.hi, g.Sf, c->mheavy + g.mhoff_f, g.mnhf, c->gFptr,
                                c->gFidx, gFw, y, rin, y);
  const TailProgram& TP = P.tail;
  if (c->prof_sweeps) (void)hipEventRecord(c->ev[5], c->stream);
  if (TP.NT > 0) {
    hipLaunchKernelGGL((k_mtail_ext<K, T>), dim3(grid_for(TP.NT * K, kBlock)), dim3(kBlock), 0,
                       c->stream, (int)TP.T0, (int)TP.NT, c->Eptr, c->Eidx, Ew, rin, y, ext,
                       m.flags);
    if (c->prof_sweeps) (void)hipEventRecord(c->ev[6], c->stream);
    const DenseBlock& DB = P.dense;
    if (DB.K == 0) {
      LAPGPU_DISPATCH_E(TP.elog, hipLaunchKernelGGL, (k_mtail<K, T, S, E_>), dim3(K),
                        dim3(kTailThreads), 0, c->stream, (int)TP.T0, (int)TP.NT, ext, sv, z,
                        c->fsteps, (int)TP.fsteps.size(), fslots, c->bsteps,
                        (int)TP.bsteps.size(), bslots, m.st, m.flags);
    } else {
      const MDenseArgs<T> da = mdense_args<T>(c);
      const int NT1 = (int)TP.NT1;
      T* u = NT1 > 0 ? da.dvec : ext;
      if (NT1 > 0)
        LAPGPU_DISPATCH_E(TP.elog, hipLaunchKernelGGL, (k_mtail_fwd<K, T, S, E_>), dim3(K),
                          dim3(kTailThreads), 0, c->stream, (int)TP.T0, (int)TP.NT, NT1, ext, sv,
                          z, da.dvec, c->fsteps, (int)TP.fsteps.size(), fslots, m.st, m.flags);
      const int tiles = (int)((DB.K + kMDenseRows - 1) / kMDenseRows);
      const int nsplit = dense_splits(c);
      const long len = (long)((c->ldd / nsplit + kMDenseSlab - 1) / kMDenseSlab * kMDenseSlab);
      if (c->prof_sweeps) (void)hipEventRecord(c->ev[8], c->stream);
      hipLaunchKernelGGL((k_mdense<K, T>), dim3(tiles, nsplit), dim3(kBlock), 0, c->stream,
                         (int)DB.K, (long)c->ldd, da.D, u, da.part, len, m.flags);
      hipLaunchKernelGGL((k_mdense_sum<K, T>), dim3(kMRed), dim3(kBlock), 0, c->stream, (int)DB.K,
                         nsplit, da.part, z + (size_t)DB.T1 * K, m.flags);
      if (c->prof_sweeps) (void)hipEventRecord(c->ev[9], c->stream);
      if (NT1 > 0)
        LAPGPU_DISPATCH_E(TP.elog, hipLaunchKernelGGL, (k_mtail_bwd<K, T, S, E_>), dim3(K),
                          dim3(kTailThreads), 0, c->stream, (int)TP.T0, (int)TP.NT, NT1, z,
                          c->bsteps, (int)TP.bsteps.size(), bslots, m.st, m.flags);
    }
  } else if (c->prof_sweeps) {
    (void)hipEventRecord(c->ev[6], c->stream);
  }
  if (c->prof_sweeps) (void)hipEventRecord(c->ev[7], c->stream);
  for (auto it = c->groups.rbegin(); it != c->groups.rend(); ++it)
    launch_mgroup<K, T, T>(c, it->lo, it->hi, it->Sb, c->mheavy + it->mhoff_b, it->mnhb,
                           c->gBptr, c->gBidx, gBw, z, y, z);
}

template <int K>
void msweeps(lapgpu_ctx* c, const double* rin) {
  auto& m = c->m;
  if (c->precision == LAPGPU_FP32_INNER)
    msweeps_t<K, float, TailSlotF>(c, rin, m.y32, m.z32, m.ext32, c->gFw32, c->gBw32, c->s32,
                                   c->Ew32, c->fslots32, c->bslots32);
  else
    msweeps_t<K, double, TailSlotD>(c, rin, m.y, m.z, m.ext, c->gFw, c->gBw, c->s, c->Ew,
                                    c->fslots, c->bslots);
}

// mean removal partials and (yF - mu).r after the sweeps
template <int K>
void mfinish_precond(lapgpu_ctx* c, const double* rin) {
  auto& m = c->m;
  const int n = (int)c->n;
  if (c->precision == LAPGPU_FP32_INNER) {
    hipLaunchKernelGGL((k_mvsum<K, float>), dim3(kMRed), dim3(kBlock), 0, c->stream, n, m.z32,
                       m.part, m.flags);
    hipLaunchKernelGGL(k_mscalar_mu<K>, dim3(1), dim3(kScalarThreads), 0, c->stream, n, m.part,
                       m.st, m.flags);
    hipLaunchKernelGGL((k_mzdot<K, float>), dim3(kMRed), dim3(kBlock), 0, c->stream, n, m.z32,
                       rin, m.p, m.part, m.st, m.flags);
  } else {
    hipLaunchKernelGGL((k_mvsum<K, double>), dim3(kMRed), dim3(kBlock), 0, c->stream, n, m.z,
                       m.part, m.flags);
    hipLaunchKernelGGL(k_mscalar_mu<K>, dim3(1), dim3(kScalarThreads), 0, c->stream, n, m.part,
                       m.st, m.flags);
    hipLaunchKernelGGL((k_mzdot<K, double>), dim3(kMRed), dim3(kBlock), 0, c->stream, n, m.z,
                       rin, m.p, m.part, m.st, m.flags);
  }
}

// p = yF - mu (FP64 and, in the FP32-inner mode, the FP32 copy) or, with
// `update`, p = (yF - mu) + beta p
template <int K>
void mp_step(lapgpu_ctx* c, bool update) {
  auto& m = c->m;
  const int n = (int)c->n;
  const bool f32 = c->precision == LAPGPU_FP32_INNER;
  if (update) {
    if (f32)
      hipLaunchKernelGGL((k_mpupdate<K, float>), dim3(kMRed), dim3(kBlock), 0, c->stream, n,
                         m.z32, m.p, m.pf, m.st, m.flags);
    else
      hipLaunchKernelGGL((k_mpupdate<K, double>), dim3(kMRed), dim3(kBlock), 0, c->stream, n,
                         m.z, m.p, (float*)nullptr, m.st, m.flags);
  } else {
    if (f32)
      hipLaunchKernelGGL((k_mpinit<K, float>), dim3(kMRed), dim3(kBlock), 0, c->stream, n, m.z32,
                         m.p, m.pf, m.st, m.flags);
    else
      hipLaunchKernelGGL((k_mpinit<K, double>), dim3(kMRed), dim3(kBlock), 0, c->stream, n, m.z,
                         m.p, (float*)nullptr, m.st, m.flags);
  }
}

// one PCG iteration for K right-hand sides; `replace`: FP32-mode iteration
// that ends with an FP64 residual replacement
template <int K>
void enqueue_miter(lapgpu_ctx* c, bool prof, bool replace) {
  auto& m = c->m;
  const int n = (int)c->n;
  const bool f32 = c->precision == LAPGPU_FP32_INNER;
  if (prof) (void)hipEventRecord(c->ev[0], c->stream);
  if (f32)
    hipLaunchKernelGGL(k_mspmv_mixed<K>, dim3(kMRed), dim3(kBlock), 0, c->stream, n, c->Lptr,
                       c->Lidx, c->Lval32, m.pf, c->Lval, m.p, m.q, m.part, m.st, m.flags);
  else
    hipLaunchKernelGGL(k_mspmv<K>, dim3(kMRed), dim3(kBlock), 0, c->stream, n, c->Lptr, c->Lidx,
                       c->Lval, m.p, m.q, m.part, m.flags);
  if (prof) (void)hipEventRecord(c->ev[1], c->stream);
  hipLaunchKernelGGL(k_mscalar_al<K>, dim3(1), dim3(kScalarThreads), 0, c->stream, m.part, m.st,
                     m.flags);
  hipLaunchKernelGGL(k_mupdate<K>, dim3(kMRed), dim3(kBlock), 0, c->stream, n, m.x, m.bestx,
                     m.p, m.q, m.r, m.part, m.st, m.flags);
  hipLaunchKernelGGL(k_mscalar_nr<K>, dim3(1), dim3(kScalarThreads), 0, c->stream, m.part, m.st,
                     m.flags);
  if (f32) {
    hipLaunchKernelGGL(k_mresid<K>, dim3(kMRed), dim3(kBlock), 0, c->stream, n, replace ? 1 : 0,
                       c->Lptr, c->Lidx, c->Lval, m.x, m.r0, m.r, m.part, m.st, m.flags);
    hipLaunchKernelGGL(k_mscalar_resid<K>, dim3(1), dim3(kScalarThreads), 0, c->stream,
                       replace ? 1 : 0, m.part, m.st, m.flags);
  }
  if (prof) (void)hipEventRecord(c->ev[2], c->stream);
  c->prof_sweeps = prof;
  msweeps<K>(c, m.r);
  c->prof_sweeps = false;
  if (prof) (void)hipEventRecord(c->ev[3], c->stream);
  mfinish_precond<K>(c, m.r);
  hipLaunchKernelGGL(k_mscalar_rho<K>, dim3(1), dim3(kScalarThreads), 0, c->stream, m.part,
                     m.st, m.flags);
  mp_step<K>(c, true);
  if (prof) (void)hipEventRecord(c->ev[4], c->stream);
}

// columns [0, k) of d_B (k <= K), results to d_X
template <int K>
int solve_multi_chunk(lapgpu_ctx* c, const double* d_B, double* d_X, int k, double tol,
                      int64_t maxits, double maxtime, int stag_test, int flags,
                      int64_t* its_out, double* relres_out, int* stop_out, double* acc) {
  auto& m = c->m;
  const bool prof = (flags & LAPGPU_PROFILE) != 0;
  const bool f32 = c->precision == LAPGPU_FP32_INNER;
  const int n = (int)c->n;
  int rc = multi_prepare(c, K);
  if (rc) return rc;
  for (int j = 0; j < K; ++j)
    fill_state(c, m.hst[j], tol, maxits, maxtime, stag_test, flags, f32);
  HIPC(c, hipMemcpyAsync(m.st, m.hst, sizeof(DevState) * K, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemsetAsync(m.flags, 0, 4 * sizeof(int), c->stream));

  // setup: r = b - mean(b), z = pre(r), p = z, rho = r.z (q holds b)
  hipLaunchKernelGGL(k_mgather<K>, dim3(kMRed), dim3(kBlock), 0, c->stream, n, k, c->perm, d_B,
                     m.q, m.part);
  hipLaunchKernelGGL(k_mscalar_mean<K>, dim3(1), dim3(kScalarThreads), 0, c->stream, n, m.part,
                     m.st);
  hipLaunchKernelGGL(k_mcenter<K>, dim3(kMRed), dim3(kBlock), 0, c->stream, n, m.q, m.r, m.x,
                     m.bestx, m.part, m.st);
  hipLaunchKernelGGL(k_mscalar_nb<K>, dim3(1), dim3(kScalarThreads), 0, c->stream, m.part, m.st,
                     m.flags);
  if (f32)
    HIPC(c, hipMemcpyAsync(m.r0, m.r, sizeof(double) * n * K, hipMemcpyDeviceToDevice,
                           c->stream));
  msweeps<K>(c, m.r);
  mfinish_precond<K>(c, m.r);
  hipLaunchKernelGGL(k_mscalar_rho0<K>, dim3(1), dim3(kScalarThreads), 0, c->stream, m.part,
                     m.st, m.flags);
  mp_step<K>(c, false);
  HIPC(c, hipGetLastError());
  HIPC(c, hipMemcpyAsync(&c->hflags[0], m.flags, sizeof(int32_t), hipMemcpyDeviceToHost,
                         c->stream));
  HIPC(c, sync_stream(c));
  bool done = c->hflags[0] != 0;

  const int rep_every = f32 ? c->replace_every : 0;
  const bool captured = !prof && c->use_graph && !(flags & LAPGPU_EAGER);
  if (captured && m.exec_K != K) {
    drop_graphs(c);
    m.exec_K = K;
  }
  if (captured) {
    for (int v = 0; v < (f32 ? 2 : 1); ++v) {
      if (m.exec[v]) continue;
      LAPGPU_LOCK();
      hipGraph_t g = nullptr;
      HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      enqueue_miter<K>(c, false, v == 1);
      HIPC(c, hipStreamEndCapture(c->stream, &g));
      hipError_t ie = hipGraphInstantiate(&m.exec[v], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPC(c, ie);
    }
  }
  int64_t enq = 0, polled = 0;
  while (!done) {
    const bool replace = rep_every > 0 && ((enq + 1) % rep_every == 0);
    hipGraphExec_t ex = m.exec[replace ? 1 : 0];
    if (captured && ex)
      HIPC(c, hipGraphLaunch(ex, c->stream));
    else
      enqueue_miter<K>(c, prof, replace);
    const int slot = (int)(enq % lapgpu_ctx::kRing);
    HIPC(c, hipMemcpyAsync(&c->hflags[slot], m.flags, sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
    HIPC(c, hipEventRecord(c->ring_ev[slot], c->stream));
    ++enq;
    if (prof) {
      HIPC(c, hipEventSynchronize(c->ev[4]));
      float a = 0, b2 = 0, d = 0, e = 0;
      (void)hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
      (void)hipEventElapsedTime(&b2, c->ev[2], c->ev[3]);
      (void)hipEventElapsedTime(&d, c->ev[1], c->ev[2]);
      (void)hipEventElapsedTime(&e, c->ev[3], c->ev[4]);
      acc[0] += a;
      acc[1] += b2;
      acc[2] += d + e;
      acc[3] += 1;
      float f1 = 0, f2 = 0, f3 = 0, f4 = 0;
      (void)hipEventElapsedTime(&f1, c->ev[2], c->ev[5]);
      (void)hipEventElapsedTime(&f2, c->ev[5], c->ev[6]);
      (void)hipEventElapsedTime(&f3, c->ev[6], c->ev[7]);
      (void)hipEventElapsedTime(&f4, c->ev[7], c->ev[3]);
      acc[4] += f1;
      acc[5] += f2;
      acc[6] += f3;
      acc[7] += f4;
      if (c->plan.dense.K > 0) {
        float g1 = 0, g2 = 0, g3 = 0;
        (void)hipEventElapsedTime(&g1, c->ev[6], c->ev[8]);
        (void)hipEventElapsedTime(&g2, c->ev[8], c->ev[9]);
        (void)hipEventElapsedTime(&g3, c->ev[9], c->ev[7]);
        acc[8] += g1;
        acc[9] += g2;
        acc[10] += g3;
      }
    }
    while (!done && polled < enq - 1) {
      const int ps = (int)(polled % lapgpu_ctx::kRing);
      HIPC(c, wait_event(c, c->ring_ev[ps]));
      ++polled;
      if (c->hflags[ps]) done = true;
    }
    // cancelled columns stop after the iterations already enqueued; the
    // others run on (all stopped: the next iteration raises flags[0])
    if (!done) {
      if (const uint64_t mk = take_cancel(c, c->chunk0, k))
        hipLaunchKernelGGL(k_cancel, dim3(1), dim3(kWave), 0, c->stream, m.st, k, mk);
    }
  }
  // the true residual reads B, which the final copy may overwrite (X = B)
  hipLaunchKernelGGL(k_mtrue_resid<K>, dim3(kMRed), dim3(kBlock), 0, c->stream, n, k, c->perm,
                     c->Lptr, c->Lidx, c->Lval, m.x, m.bestx, d_B, m.part, m.st);
  hipLaunchKernelGGL(k_mscalar_true<K>, dim3(1), dim3(kScalarThreads), 0, c->stream, m.part,
                     m.st);
  hipLaunchKernelGGL(k_mfinal<K>, dim3(kMRed), dim3(kBlock), 0, c->stream, n, k, c->perm, m.x,
                     m.bestx, d_X, m.st);
  HIPC(c, hipMemcpyAsync(m.hst, m.st, sizeof(DevState) * K, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, sync_stream(c));
  HIPC(c, hipGetLastError());
  for (int j = 0; j < k; ++j) {
    const DevState& h = m.hst[j];
    if (its_out) its_out[j] = h.itcnt;
    if (relres_out) relres_out[j] = h.nr;
    if (stop_out) stop_out[j] = h.stop_reason;
    if ((size_t)(c->chunk0 + j) < c->last_true.size()) c->last_true[c->chunk0 + j] = h.true_nr;
    if (h.stop_reason == 3) printf("PCG Stopped by stagnation test %d\n", stag_test);
    if ((flags & LAPGPU_VERBOSE) && h.stop_reason != 9)
      printf("rhs %d: PCG stopped after %lld iterations with relative error %.17g (%s).\n", j,
             (long long)h.itcnt, h.nr, h.stop_reason == 1 ? "maxits" : stop_message(h.stop_reason));
  }
  fflush(stdout);
  return 0;
}

int solve_multi_device(lapgpu_ctx* c, const double* d_B, double* d_X, int64_t k, double tol,
                       int64_t maxits, double maxtime, int stag_test, int flags,
                       int64_t* its_out, double* relres_out, int* stop_out) {
  if (!c->has_F) return set_err(c, -2, "solve_multi: upload the Laplacian and LDLinv first");
  const auto t_start = std::chrono::steady_clock::now();
  const int64_t n = c->n;
  // spmv, ldl, vec, iterations; forward groups, tail inputs, tail, backward
  // groups; sparse tail forward, dense product, sparse tail backward
  double acc[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  int64_t maxit = 0;
  for (int64_t c0 = 0; c0 < k; c0 += kMultiMax) {
    const int kc = (int)std::min<int64_t>(kMultiMax, k - c0);
    int K = 8;
    while (K < kc) K *= 2;
    int rc = 0;
    c->chunk0 = c->col_base + c0;
    LAPGPU_DISPATCH_K(K, rc = solve_multi_chunk<K_>, c, d_B + c0 * n, d_X + c0 * n, kc, tol,
                      maxits, maxtime, stag_test, flags, its_out ? its_out + c0 : nullptr,
                      relres_out ? relres_out + c0 : nullptr, stop_out ? stop_out + c0 : nullptr,
                      acc);
    if (rc) return rc;
    for (int j = 0; j < kc; ++j) maxit = std::max<int64_t>(maxit, c->m.hst[j].itcnt);
  }
  if ((flags & LAPGPU_PROFILE) && acc[3] > 0) {
    c->t_spmv = acc[0] / acc[3];
    c->t_ldl = acc[1] / acc[3];
    c->t_vec = acc[2] / acc[3];
    c->t_fgroups = acc[4] / acc[3];
    c->t_tail_ext = acc[5] / acc[3];
    c->t_tail = acc[6] / acc[3];
    c->t_bgroups = acc[7] / acc[3];
    c->t_tail_fwd = acc[8] / acc[3];
    c->t_dense = acc[9] / acc[3];
    c->t_tail_bwd = acc[10] / acc[3];
  }
  c->t_total =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() * 1e3;
  c->t_its = maxit;
  c->err.clear();
  return 0;
}

// ---- autotuner: lanes per row (single-RHS) and slices per row (multi-RHS)
// of every level group and sweep direction, and optionally the level-group
// budget, chosen by timing the candidates on this graph.  The sweep inputs
// are whatever the workspace holds (only timings are used).

constexpr int kTuneReps = 4;

// mean device time (ms) of one call of `launch` over kTuneReps calls, after
// one untimed call
template <typename F>
float tune_time(lapgpu_ctx* c, F&& launch) {
  launch();
  (void)hipEventRecord(c->ev[0], c->stream);
  for (int r = 0; r < kTuneReps; ++r) launch();
  (void)hipEventRecord(c->ev[1], c->stream);
  (void)hipEventSynchronize(c->ev[1]);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
  return ms / (float)kTuneReps;
}

// rows of [lo, hi) longer than 8 G entries (the wave-per-row rows at G)
void heavy_rows(const std::vector<int32_t>& ptr, int lo, int hi, int G, std::vector<int32_t>& out) {
  for (int v = lo; v < hi; ++v)
    if (ptr[v + 1] - ptr[v] > 8 * G) out.push_back(v);
}

template <typename T>
int tune_lanes_t(lapgpu_ctx* c, const double* rin, T* y, T* z, const T* gFw, const T* gBw,
                 double* before_ms, double* after_ms) {
  const LevelBlocks& B = c->plan.blocks;
  constexpr int kCand[4] = {4, 8, 16, 32};
  // the heavy rows of every candidate, per group and direction
  std::vector<int32_t> hv[4];
  std::vector<std::array<int, 4>> off(c->groups.size());  // f, nf, b, nb
  std::vector<std::array<int, 4>> offc[4];
  for (int ci = 0; ci < 4; ++ci) {
    offc[ci].resize(c->groups.size());
    for (size_t g = 0; g < c->groups.size(); ++g) {
      const auto& gl = c->groups[g];
      const int f0 = (int)hv[ci].size();
      heavy_rows(B.fptr, gl.lo, gl.hi, kCand[ci], hv[ci]);
      const int b0 = (int)hv[ci].size();
      heavy_rows(B.bptr, gl.lo, gl.hi, kCand[ci], hv[ci]);
      offc[ci][g] = {f0, b0 - f0, b0, (int)hv[ci].size() - b0};
    }
  }
  int* dh[4] = {nullptr, nullptr, nullptr, nullptr};
  int rc = 0;
  for (int ci = 0; ci < 4 && !rc; ++ci) rc = dev_upload(c, &dh[ci], hv[ci]);
  double t_before = 0.0, t_after = 0.0;
  for (size_t g = 0; g < c->groups.size() && !rc; ++g) {
    auto& gl = c->groups[g];
    for (int dir = 0; dir < 2; ++dir) {
      const int cur = dir == 0 ? gl.Gf : gl.Gb;
      float best = 0.0f;
      int bestG = cur;
      for (int ci = 0; ci < 4; ++ci) {
        const int G = kCand[ci];
        const int* h = dh[ci] + offc[ci][g][dir * 2];
        const int nh = offc[ci][g][dir * 2 + 1];
        const float t = tune_time(c, [&] {
          if (dir == 0) {
            LAPGPU_DISPATCH_G(G, (launch_group<G_, T, double>), c, gl.lo, gl.hi, h, nh,
                              c->gFptr, c->gFidx, gFw, y, rin, y);
          } else {
            LAPGPU_DISPATCH_G(G, (launch_group<G_, T, T>), c, gl.lo, gl.hi, h, nh, c->gBptr,
                              c->gBidx, gBw, z, y, z);
          }
        });
        if (G == cur) t_before += t;
        if (ci == 0 || t < best) {
          best = t;
          bestG = G;
        }
      }
      t_after += best;
      (dir == 0 ? gl.Gf : gl.Gb) = bestG;
    }
  }
  for (int ci = 0; ci < 4; ++ci) dev_free(c, dh[ci]);
  if (rc) return rc;
  // the final heavy list, in the layout upload_groups makes
  std::vector<int32_t> heavy;
  for (auto& gl : c->groups) {
    gl.hoff_f = (int)heavy.size();
    heavy_rows(B.fptr, gl.lo, gl.hi, gl.Gf, heavy);
    gl.nhf = (int)heavy.size() - gl.hoff_f;
    gl.hoff_b = (int)heavy.size();
    heavy_rows(B.bptr, gl.lo, gl.hi, gl.Gb, heavy);
    gl.nhb = (int)heavy.size() - gl.hoff_b;
  }
  if ((rc = dev_upload(c, &c->heavy, heavy))) return rc;
  if ((rc = build_chain(c))) return rc;
  *before_ms += t_before;
  *after_ms += t_after;
  return 0;
}

template <int K, typename T>
int tune_slices_t(lapgpu_ctx* c, T* y, T* z, const T* gFw, const T* gBw, double* before_ms,
                  double* after_ms) {
  auto& m = c->m;
  HIPC(c, hipMemsetAsync(m.flags, 0, 4 * sizeof(int), c->stream));  // no column stopped
  double t_before = 0.0, t_after = 0.0;
  for (auto& gl : c->groups) {
    for (int dir = 0; dir < 2; ++dir) {
      const int cur = dir == 0 ? gl.Sf : gl.Sb;
      float best = 0.0f;
      int bestS = cur;
      for (int S = 1; S <= 8 && S * K <= kWave; S *= 2) {
        const float t = tune_time(c, [&] {
          if (dir == 0)
            launch_mgroup<K, T, double>(c, gl.lo, gl.hi, S, c->mheavy + gl.mhoff_f, gl.mnhf,
                                        c->gFptr, c->gFidx, gFw, y, m.r, y);
          else
            launch_mgroup<K, T, T>(c, gl.lo, gl.hi, S, c->mheavy + gl.mhoff_b, gl.mnhb,
                                   c->gBptr, c->gBidx, gBw, z, y, z);
        });
        if (S == std::min(cur, kWave / K)) t_before += t;
        if (S == 1 || t < best) {
          best = t;
          bestS = S;
        }
      }
      t_after += best;
      (dir == 0 ? gl.Sf : gl.Sb) = bestS;
    }
  }
  *before_ms += t_before;
  *after_ms += t_after;
  return 0;
}

template <int K>
int tune_slices(lapgpu_ctx* c, double* before_ms, double* after_ms) {
  int rc = multi_prepare(c, K);
  if (rc) return rc;
  auto& m = c->m;
  if (c->precision == LAPGPU_FP32_INNER)
    return tune_slices_t<K, float>(c, m.y32, m.z32, c->gFw32, c->gBw32, before_ms, after_ms);
  return tune_slices_t<K, double>(c, m.y, m.z, c->gFw, c->gBw, before_ms, after_ms);
}

int tune_lanes(lapgpu_ctx* c, double* before_ms, double* after_ms) {
  if (c->precision == LAPGPU_FP32_INNER)
    return tune_lanes_t<float>(c, c->r, c->y32, c->z32, c->gFw32, c->gBw32, before_ms, after_ms);
  return tune_lanes_t<double>(c, c->r, c->y, c->zF, c->gFw, c->gBw, before_ms, after_ms);
}

// device time (ms) of one LDLinv apply (sweeps only) with the current schedule
float apply_time(lapgpu_ctx* c) { return tune_time(c, [&] { sweeps(c, c->r); }); }

template <int K>
float mapply_time(lapgpu_ctx* c) {
  auto& m = c->m;
  // no column stopped (a solve sets these states again when it starts)
  (void)hipMemsetAsync(m.flags, 0, 4 * sizeof(int), c->stream);
  (void)hipMemsetAsync(m.st, 0, sizeof(DevState) * K, c->stream);
  return tune_time(c, [&] { msweeps<K>(c, m.r); });
}

// rebuild the level groups with the context's budget (captured iterations
// and the FP32 group weights follow)
int regroup(lapgpu_ctx* c) {
  drop_graphs(c);
  int rc = upload_groups(c, false);
  if (rc) return rc;
  if (c->f32_ready) {
    if ((rc = dev_upload(c, &c->gFw32, narrow<float>(c->plan.blocks.fw)))) return rc;
    if ((rc = dev_upload(c, &c->gBw32, narrow<float>(c->plan.blocks.bw)))) return rc;
  }
  return 0;
}

}  // namespace

extern "C" {

int lapgpu_create(lapgpu_ctx** out, int device) {
  if (!out) return set_err(nullptr, -1, "lapgpu_create: null output");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return set_err(nullptr, -10, "lapgpu_create: no HIP device visible");
  if (device < 0 || device >= count)
    return set_err(nullptr, -1, "lapgpu_create: device index out of range");
  LAPGPU_LOCK();
  lapgpu_ctx* c = new lapgpu_ctx();
  c->device = device;
  auto fail = [&](const char* what, hipError_t e) {
    g_create_err = std::string(what) + ": " + hipGetErrorString(e);
    delete c;
    return -10;
  };
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return fail("hipSetDevice", e);
  if ((e = lapgpu::pool_setup(device)) != hipSuccess) return fail("hipMemPoolSetAttribute", e);
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
    return fail("hipStreamCreate", e);
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
    c->wallclock_khz = (double)khz;
  if ((e = hipMallocAsync((void**)&c->part, sizeof(double) * kNumSlots * kRedBlocks, c->stream)) !=
      hipSuccess)
    return fail("hipMallocAsync", e);
  if ((e = hipMallocAsync((void**)&c->st, sizeof(DevState), c->stream)) != hipSuccess)
    return fail("hipMallocAsync", e);
  if ((e = lapgpu::pinned_get((void**)&c->hst, sizeof(DevState))) != hipSuccess)
    return fail("hipHostMalloc", e);
  if ((e = lapgpu::pinned_get((void**)&c->hflags, sizeof(int32_t) * lapgpu_ctx::kRing)) !=
      hipSuccess)
    return fail("hipHostMalloc", e);
  for (auto& ev : c->ring_ev)
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
      return fail("hipEventCreate", e);
  for (auto& ev : c->ev)
    if ((e = hipEventCreate(&ev)) != hipSuccess) return fail("hipEventCreate", e);
  if ((e = lapgpu::pinned_get((void**)&c->hchain, sizeof(int32_t))) != hipSuccess)
    return fail("hipHostMalloc", e);
  c->hchain[0] = 0;
  if (const char* ch = getenv("LAPGPU_CHAIN")) c->use_chain = atoi(ch) != 0;
  if (const char* cn = getenv("LAPGPU_CHAIN_NARROW")) c->chain_narrow = atoi(cn);
  if (const char* cm = getenv("LAPGPU_CHAIN_MIN_RUN")) c->chain_min_run = atoi(cm);
  if (const char* cp = getenv("LAPGPU_CHAIN_POLL_MAX")) c->chain_poll_max = std::max(1, atoi(cp));
  if (const char* mk = getenv("LAPGPU_MGROUP_MERGE_K")) c->mgroup_merge_k = atoi(mk);
  if (const char* tw = getenv("LAPGPU_TAIL_STEPS")) c->max_tail_steps = atoll(tw);
  if (const char* ug = getenv("LAPGPU_NO_GRAPH")) c->use_graph = atoi(ug) == 0;
  if (const char* sw = getenv("LAPGPU_SWITCH_TOL")) c->switch_tol = atof(sw);
  if (const char* bd = getenv("LAPGPU_BLOCK_DEPTH")) c->block_depth = atoi(bd);
  if (const char* bg = getenv("LAPGPU_BLOCK_GROWTH")) c->block_growth = atof(bg);
  if (const char* be = getenv("LAPGPU_BLOCK_EXTRA")) c->block_extra = atoll(be);
  if (const char* dm = getenv("LAPGPU_DENSE_MAX"))
    c->dense_max = std::max<int64_t>(0, std::min<int64_t>(kTailCap, atoll(dm)));
  *out = c;
  return 0;
}

void lapgpu_destroy(lapgpu_ctx* c) {
  if (!c) return;
  LAPGPU_LOCK();
  (void)hipSetDevice(c->device);
  if (c->stream) {
    (void)hipStreamSynchronize(c->stream);
    free_graph(c);
    dev_free(c, c->part);
    dev_free(c, c->st);
    (void)hipStreamSynchronize(c->stream);
  }
  lapgpu::pinned_put(c->hst, sizeof(DevState));
  lapgpu::pinned_put(c->hchain, sizeof(int32_t));
  lapgpu::pinned_put(c->hflags, sizeof(int32_t) * lapgpu_ctx::kRing);
  for (auto& ev : c->ring_ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& ev : c->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (c->sync_ev) (void)hipEventDestroy(c->sync_ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int lapgpu_autotune(lapgpu_ctx* c, int64_t k, int flags, double* times_ms) {
  if (!c) return -1;
  if (!c->has_F) return set_err(c, -2, "autotune: upload the Laplacian and LDLinv first");
  if (k < 0 || k > kMultiMax) return set_err(c, -1, "autotune: k must be in [0, 64]");
  HIPC(c, hipSetDevice(c->device));
  int K = 0;
  if (k > 1) {
    K = 8;
    while (K < k) K *= 2;
  }
  int rc = 0;
  double t[4] = {0, 0, 0, 0};
  // multi-RHS workspace first (the apply timings below need it)
  if (K) LAPGPU_DISPATCH_K(K, rc = multi_prepare, c, K_);
  if (rc) return rc;
  // a running state (a solve sets it again when it starts)
  HIPC(c, hipMemsetAsync(c->st, 0, sizeof(DevState), c->stream));
  t[0] = apply_time(c);
  if (K) LAPGPU_DISPATCH_K(K, t[2] = mapply_time<K_>, c);
  if (flags & LAPGPU_TUNE_GROUPS) {
    // the level-group budget: each candidate regroups and times an apply
    // (of K right-hand sides when K > 0): the context's depth and growth
    // with three extra-entry limits, and two small budgets whose many
    // narrow groups the chained runs take (the fastest single-RHS
    // schedules of expanders)
    struct Budget {
      int depth;
      double growth;
      int64_t extra;
    };
    const Budget cand[5] = {{c->block_depth, c->block_growth, 65536},
                            {c->block_depth, c->block_growth, 196608},
                            {c->block_depth, c->block_growth, 524288},
                            {16, 4.0, 65536},
                            {8, 2.0, 16384}};
    float best = 0.0f;
    Budget pick = {c->block_depth, c->block_growth, c->block_extra};
    for (int i = 0; i < 5; ++i) {
      c->block_depth = cand[i].depth;
      c->block_growth = cand[i].growth;
      c->block_extra = cand[i].extra;
      if ((rc = regroup(c))) return rc;
      float tt = apply_time(c);
      if (K) LAPGPU_DISPATCH_K(K, tt = mapply_time<K_>, c);
      if (i == 0 || tt < best) {
        best = tt;
        pick = cand[i];
      }
    }
    c->block_depth = pick.depth;
    c->block_growth = pick.growth;
    c->block_extra = pick.extra;
    if ((rc = regroup(c))) return rc;
  }
  double b = 0.0, a = 0.0;
  if ((rc = tune_lanes(c, &b, &a))) return rc;
  if (K) {
    LAPGPU_DISPATCH_K(K, rc = tune_slices<K_>, c, &b, &a);
    if (rc) return rc;
  }
  drop_graphs(c);  // the next solve captures the tuned schedule
  t[1] = apply_time(c);
  if (K) LAPGPU_DISPATCH_K(K, t[3] = mapply_time<K_>, c);
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipGetLastError());
  if (times_ms)
    for (int i = 0; i < 4; ++i) times_ms[i] = t[i];
  c->err.clear();
  return 0;
}

const char* lapgpu_last_error(lapgpu_ctx* c) {
  return c ? c->err.c_str() : g_create_err.c_str();
}

int lapgpu_upload_csc(lapgpu_ctx* c, int64_t n, const int64_t* colptr,
                      const int64_t* rowval, const double* nzval,
                      int index_base) {
  if (!c) return -1;
  HIPC(c, hipSetDevice(c->device));
  std::string err;
  HostCsr L;
  if (!csc_to_csr(n, colptr, rowval, nzval, index_base, L, err)) return set_err(c, -1, err);
  c->L = std::move(L);
  c->has_L = true;
  c->has_F = false;
  identity_plan(c);
  int rc = upload_plan(c);
  if (rc) return rc;
  c->err.clear();
  return 0;
}

int lapgpu_upload_ldli(lapgpu_ctx* c, int64_t n, int64_t nnzF,
                       const int64_t* col, const int64_t* colptr,
                       const int64_t* rowval, const double* fval,
                       const double* d, int index_base) {
  if (!c) return -1;
  if (!c->has_L || c->L.n == 0)
    return set_err(c, -2, "upload_ldli: upload the Laplacian first (a loaded plan keeps only its own)");
  if (n != c->L.n) return set_err(c, -1, "upload_ldli: n differs from the Laplacian's");
  if (index_base != 0 && index_base != 1)
    return set_err(c, -1, "upload_ldli: index_base must be 0 or 1");
  HIPC(c, hipSetDevice(c->device));
  std::string err;
  Plan P;
  DenseModel dm;
  dm.max_k = c->dense_max;
  if (!build_plan(c->L, nnzF, col, colptr, rowval, fval, d, index_base, c->max_tail_steps, P, err, dm))
    return set_err(c, -1, err);
  c->plan = std::move(P);
  int rc = upload_plan(c);
  if (rc) return rc;
  c->has_F = true;
  c->err.clear();
  return 0;
}

int lapgpu_save_plan(lapgpu_ctx* c, const char* path) {
  if (!c || !path) return -1;
  if (!c->has_F) return set_err(c, -2, "save_plan: upload the Laplacian and LDLinv (or load a plan) first");
  std::string err;
  if (!save_plan(c->plan, path, err)) return set_err(c, -1, err);
  c->err.clear();
  return 0;
}

int lapgpu_load_plan(lapgpu_ctx* c, const char* path, int64_t* n_out) {
  if (!c || !path) return -1;
  HIPC(c, hipSetDevice(c->device));
  std::string err;
  Plan P;
  if (!load_plan(path, P, err)) return set_err(c, -1, err);
  HIPC(c, hipStreamSynchronize(c->stream));
  c->plan = std::move(P);
  // la lives in the plan, permuted; the caller-order copy is not kept
  c->L = HostCsr();
  c->has_L = true;
  c->has_F = false;
  int rc = upload_plan(c, true);
  if (rc) return rc;
  c->has_F = true;
  if (n_out) *n_out = c->n;
  c->err.clear();
  return 0;
}

int lapgpu_solve_device(lapgpu_ctx* c, const double* d_b, double* d_x,
                        int64_t n, double tol, int64_t maxits, double maxtime,
                        int stag_test, int flags, int64_t* its_out,
                        double* relres_out, int* stop_out) {
  if (!c) return -1;
  if (n != c->n) return set_err(c, -1, "solve: length of b differs from n");
  if (!d_b || !d_x) return set_err(c, -1, "solve: null vector");
  HIPC(c, hipSetDevice(c->device));
  RunGuard run(c, 1);
  int rc = solve_device(c, d_b, d_x, tol, maxits, maxtime, stag_test, flags, its_out,
                        relres_out, stop_out);
  // b is still there unless x overwrote it
  if (rc == kChainRetry && d_b != d_x)
    rc = solve_device(c, d_b, d_x, tol, maxits, maxtime, stag_test, flags, its_out, relres_out,
                      stop_out);
  return rc;
}

int lapgpu_last_true_relres(lapgpu_ctx* c, double* out, int64_t k) {
  if (!c || !out) return -1;
  if (k != (int64_t)c->last_true.size())
    return set_err(c, -1, "last_true_relres: k differs from the last solve's column count (" +
                              std::to_string(c->last_true.size()) + ")");
  std::copy(c->last_true.begin(), c->last_true.end(), out);
  c->err.clear();
  return 0;
}

int lapgpu_cancel(lapgpu_ctx* c, int64_t column) {
  if (!c || column < -1) return -1;
  // c->err belongs to the solving thread: errors are return codes only
  std::lock_guard<std::mutex> g(c->cancel_mu);
  if (!c->running) return LAPGPU_IDLE;
  const int64_t k = (int64_t)c->cancel_cols.size();
  if (column < -1 || column >= k) return -1;
  if (column == -1)
    std::fill(c->cancel_cols.begin(), c->cancel_cols.end(), (char)1);
  else
    c->cancel_cols[(size_t)column] = 1;
  c->cancel_any.store(true, std::memory_order_release);
  return 0;
}

int lapgpu_solve(lapgpu_ctx* c, const double* b, double* x_out, int64_t n,
                 double tol, int64_t maxits, double maxtime, int stag_test,
                 int flags, int64_t* its_out, double* relres_out,
                 int* stop_out) {
  if (!c) return -1;
  if (n != c->n) return set_err(c, -1, "solve: length of b differs from n");
  if (!b || !x_out) return set_err(c, -1, "solve: null vector");
  if (!c->has_F) return set_err(c, -2, "solve: upload the Laplacian and LDLinv first");
  HIPC(c, hipSetDevice(c->device));
  RunGuard run(c, 1);
  int rc = kChainRetry;
  for (int attempt = 0; attempt < 2 && rc == kChainRetry; ++attempt) {
    HIPC(c, hipMemcpyAsync(c->io, b, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    rc = solve_device(c, c->io, c->io, tol, maxits, maxtime, stag_test, flags, its_out,
                      relres_out, stop_out);
  }
  if (rc) return rc;
  HIPC(c, hipMemcpyAsync(x_out, c->io, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, sync_stream(c));
  return 0;
}

int lapgpu_solve_multi_device(lapgpu_ctx* c, const double* d_B, double* d_X, int64_t n,
                              int64_t k, double tol, int64_t maxits, double maxtime,
                              int stag_test, int flags, int64_t* its_out, double* relres_out,
                              int* stop_out) {
  if (!c) return -1;
  if (n != c->n) return set_err(c, -1, "solve_multi: number of rows of B differs from n");
  if (k < 0 || (k > 0 && (!d_B || !d_X))) return set_err(c, -1, "solve_multi: bad arguments");
  if (k == 0) return 0;
  HIPC(c, hipSetDevice(c->device));
  RunGuard run(c, k);
  c->last_true.assign((size_t)k, NAN);
  c->col_base = 0;
  return solve_multi_device(c, d_B, d_X, k, tol, maxits, maxtime, stag_test, flags, its_out,
                            relres_out, stop_out);
}

int lapgpu_solve_multi(lapgpu_ctx* c, const double* B, double* X_out, int64_t n, int64_t k,
                       double tol, int64_t maxits, double maxtime, int stag_test, int flags,
                       int64_t* its_out, double* relres_out, int* stop_out) {
  if (!c) return -1;
  if (n != c->n) return set_err(c, -1, "solve_multi: number of rows of B differs from n");
  if (k < 0 || (k > 0 && (!B || !X_out))) return set_err(c, -1, "solve_multi: bad arguments");
  if (!c->has_F) return set_err(c, -2, "solve_multi: upload the Laplacian and LDLinv first");
  if (k == 0) return 0;
  HIPC(c, hipSetDevice(c->device));
  RunGuard run(c, k);
  c->last_true.assign((size_t)k, NAN);
  // staging: one chunk of up to 64 columns at a time through m.io
  for (int64_t c0 = 0; c0 < k; c0 += kMultiMax) {
    const int64_t kc = std::min<int64_t>(kMultiMax, k - c0);
    c->col_base = c0;
    int K = 8;
    while (K < kc) K *= 2;
    int rc = multi_prepare(c, K);
    if (rc) return rc;
    HIPC(c, hipMemcpyAsync(c->m.io, B + c0 * n, sizeof(double) * n * kc, hipMemcpyHostToDevice,
                           c->stream));
    rc = solve_multi_device(c, c->m.io, c->m.io, kc, tol, maxits, maxtime, stag_test, flags,
                            its_out ? its_out + c0 : nullptr,
                            relres_out ? relres_out + c0 : nullptr,
                            stop_out ? stop_out + c0 : nullptr);
    if (rc) return rc;
    HIPC(c, hipMemcpyAsync(X_out + c0 * n, c->m.io, sizeof(double) * n * kc,
                           hipMemcpyDeviceToHost, c->stream));
    HIPC(c, sync_stream(c));
  }
  c->err.clear();
  return 0;
}

int lapgpu_apply_ldli(lapgpu_ctx* c, const double* rin, double* zout, int64_t n) {
  if (!c) return -1;
  if (!c->has_F) return set_err(c, -2, "apply_ldli: upload the Laplacian and LDLinv first");
  if (n != c->n || !rin || !zout) return set_err(c, -1, "apply_ldli: bad arguments");
  HIPC(c, hipSetDevice(c->device));
  for (int attempt = 0;; ++attempt) {
    init_state(c, 0.0, 1, 0.0, 0, LAPGPU_NO_CENTER);
    HIPC(c, hipMemcpyAsync(c->st, c->hst, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->io, rin, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_gather_perm, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, (int)n,
                       c->perm, c->io, c->r, c->part);
    precond(c, c->r);
    hipLaunchKernelGGL(k_scatter_perm, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, (int)n,
                       c->perm, c->z, c->io);
    HIPC(c, hipGetLastError());
    HIPC(c, hipMemcpyAsync(zout, c->io, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (int rc = chain_fetch(c)) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    const int rc = chain_check(c);
    if (rc == 0) break;
    if (rc != kChainRetry || attempt > 0) return rc;
  }
  c->err.clear();
  return 0;
}

int lapgpu_spmv(lapgpu_ctx* c, const double* xin, double* yout, int64_t n) {
  if (!c) return -1;
  if (!c->has_L) return set_err(c, -2, "spmv: upload the Laplacian first");
  if (n != c->n || !xin || !yout) return set_err(c, -1, "spmv: bad arguments");
  HIPC(c, hipSetDevice(c->device));
  init_state(c, 0.0, 1, 0.0, 0, LAPGPU_NO_CENTER);
  HIPC(c, hipMemcpyAsync(c->st, c->hst, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(c->io, xin, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_gather_perm, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, (int)n,
                     c->perm, c->io, c->p, c->part);
  if (c->precision == LAPGPU_FP32_INNER && c->f32_ready)
    hipLaunchKernelGGL(k_copy_f, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, (int)n, c->p,
                       c->p, c->pf, c->st);
  spmv(c, c->p, c->q);
  hipLaunchKernelGGL(k_scatter_perm, dim3(kRedBlocks), dim3(kBlock), 0, c->stream, (int)n,
                     c->perm, c->q, c->io);
  HIPC(c, hipGetLastError());
  HIPC(c, hipMemcpyAsync(yout, c->io, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  c->err.clear();
  return 0;
}

int lapgpu_set_precision(lapgpu_ctx* c, int mode, int replace_every) {
  if (!c) return -1;
  if (mode != LAPGPU_FP64 && mode != LAPGPU_FP32_INNER)
    return set_err(c, -1, "set_precision: unknown mode");
  HIPC(c, hipSetDevice(c->device));
  HIPC(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->precision = mode;
  c->replace_every = replace_every > 0 ? replace_every : 20;
  if (mode == LAPGPU_FP32_INNER && c->has_L) {
    int rc = prepare_f32(c);
    if (rc) return rc;
  }
  c->err.clear();
  return 0;
}

int lapgpu_set_sweep_mode(lapgpu_ctx* c, int mode) {
  if (!c) return -1;
  if (mode != LAPGPU_SWEEP_GROUPS && mode != LAPGPU_SWEEP_CHAIN)
    return set_err(c, -1, "set_sweep_mode: mode must be LAPGPU_SWEEP_GROUPS or LAPGPU_SWEEP_CHAIN");
  const bool want = mode == LAPGPU_SWEEP_CHAIN;
  if (want != c->use_chain) {
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);  // captured iterations hold the other launches
    c->use_chain = want;
  }
  c->err.clear();
  return 0;
}

int lapgpu_chain_info(lapgpu_ctx* c, int64_t* runs, int64_t* chained_groups, int* active) {
  if (!c || !runs || !chained_groups || !active) return -1;
  int64_t g = 0;
  for (const auto& r : c->runs) g += r.second - r.first;
  *runs = (int64_t)c->runs.size();
  *chained_groups = g;
  *active = chain_on(c) ? 1 : 0;
  c->err.clear();
  return 0;
}

int lapgpu_set_wait_mode(lapgpu_ctx* c, int mode) {
  if (!c) return -1;
  if (mode != LAPGPU_WAIT_SPIN && mode != LAPGPU_WAIT_BLOCKING)
    return set_err(c, -1, "set_wait_mode: unknown mode");
  HIPC(c, hipSetDevice(c->device));
  HIPC(c, hipStreamSynchronize(c->stream));
  LAPGPU_LOCK();
  const unsigned fl = hipEventDisableTiming;
  for (auto& ev : c->ring_ev) {
    if (ev) (void)hipEventDestroy(ev);
    ev = nullptr;
    HIPC(c, hipEventCreateWithFlags(&ev, fl));
  }
  if (c->sync_ev) (void)hipEventDestroy(c->sync_ev);
  c->sync_ev = nullptr;
  HIPC(c, hipEventCreateWithFlags(&c->sync_ev, fl));
  c->blocking = mode == LAPGPU_WAIT_BLOCKING;
  c->err.clear();
  return 0;
}

int lapgpu_set_group_budget(lapgpu_ctx* c, int max_depth, double max_growth,
                            int64_t max_extra) {
  if (!c) return -1;
  if (max_depth < 1 || !(max_growth >= 1.0) || max_extra < 0)
    return set_err(c, -1, "set_group_budget: need max_depth >= 1, max_growth >= 1, max_extra >= 0");
  HIPC(c, hipSetDevice(c->device));
  c->block_depth = max_depth;
  c->block_growth = max_growth;
  c->block_extra = max_extra;
  if (c->has_F) {
    HIPC(c, hipStreamSynchronize(c->stream));
    int rc = regroup(c);
    if (rc) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
  }
  c->err.clear();
  return 0;
}

int lapgpu_set_dense_tail(lapgpu_ctx* c, int64_t max_k) {
  if (!c) return -1;
  if (max_k < 0 || max_k > kTailCap)
    return set_err(c, -1, "set_dense_tail: max_k must be in [0, " + std::to_string(kTailCap) + "]");
  c->dense_max = max_k;
  c->err.clear();
  return 0;
}

int lapgpu_dense_info(lapgpu_ctx* c, int64_t* k, int64_t* sparse_vertices, int64_t* bytes) {
  if (!c) return -1;
  if (!c->has_F) return set_err(c, -2, "dense_info: no LDLinv uploaded");
  const int64_t K = c->plan.dense.K;
  if (k) *k = K;
  if (sparse_vertices) *sparse_vertices = c->plan.tail.NT1;
  if (bytes) *bytes = K * c->ldd * (int64_t)(sizeof(double) + (c->D32 ? sizeof(float) : 0));
  return 0;
}

int lapgpu_sweep_entries(lapgpu_ctx* c, int64_t* groups_f, int64_t* groups_b, int64_t* tail_f,
                         int64_t* tail_b, int64_t* dense_k) {
  if (!c) return -1;
  if (!c->has_F) return set_err(c, -2, "sweep_entries: no LDLinv uploaded");
  const Plan& P = c->plan;
  const int64_t T0 = P.tail.T0, n = P.n;
  if (groups_f) *groups_f = P.Fptr[T0];
  if (groups_b) *groups_b = P.Bptr[T0];
  if (tail_f) *tail_f = P.Fptr[n] - P.Fptr[T0];
  if (tail_b) *tail_b = P.Bptr[n] - P.Bptr[T0];
  if (dense_k) *dense_k = P.dense.K;
  return 0;
}

int lapgpu_factor_info(lapgpu_ctx* c, int64_t* nlevels, int64_t* nwide,
                       int64_t* tail_start, int64_t* nnzF) {
  if (!c) return -1;
  if (!c->has_F) return set_err(c, -2, "factor_info: no LDLinv uploaded");
  if (nlevels) *nlevels = c->plan.nlev;
  if (nwide) *nwide = (int64_t)c->groups.size();
  if (tail_start) *tail_start = c->plan.tail.T0;
  if (nnzF) *nnzF = c->plan.nnzF;
  return 0;
}

int lapgpu_group_info(lapgpu_ctx* c, int64_t* ngroups, int64_t* rows, int64_t* fentries,
                      int64_t* bentries) {
  if (!c) return -1;
  if (!c->has_F) return set_err(c, -2, "group_info: no LDLinv uploaded");
  const LevelBlocks& B = c->plan.blocks;
  if (ngroups) *ngroups = (int64_t)c->groups.size();
  if (rows) *rows = c->plan.tail.T0;
  if (fentries) *fentries = B.fptr.empty() ? 0 : (int64_t)B.fptr.back();
  if (bentries) *bentries = B.bptr.empty() ? 0 : (int64_t)B.bptr.back();
  return 0;
}

int lapgpu_tail_info(lapgpu_ctx* c, int64_t* nt, int64_t* nsteps, int64_t* nslots,
                     int64_t* nedges_in) {
  if (!c) return -1;
  if (!c->has_F) return set_err(c, -2, "tail_info: no LDLinv uploaded");
  const TailProgram& TP = c->plan.tail;
  if (nt) *nt = TP.NT;
  if (nsteps) *nsteps = (int64_t)(TP.fsteps.size() + TP.bsteps.size());
  if (nslots) *nslots = (int64_t)(TP.fslots.size() + TP.bslots.size());
  if (nedges_in) *nedges_in = (int64_t)TP.Eidx.size();
  return 0;
}

int lapgpu_last_sweep_timings(lapgpu_ctx* c, double* fgroups_ms, double* tail_ext_ms,
                              double* tail_ms, double* bgroups_ms) {
  if (!c) return -1;
  if (fgroups_ms) *fgroups_ms = c->t_fgroups;
  if (tail_ext_ms) *tail_ext_ms = c->t_tail_ext;
  if (tail_ms) *tail_ms = c->t_tail;
  if (bgroups_ms) *bgroups_ms = c->t_bgroups;
  return 0;
}

int lapgpu_last_dense_timings(lapgpu_ctx* c, double* tail_fwd_ms, double* dense_ms,
                              double* tail_bwd_ms) {
  if (!c) return -1;
  if (tail_fwd_ms) *tail_fwd_ms = c->t_tail_fwd;
  if (dense_ms) *dense_ms = c->t_dense;
  if (tail_bwd_ms) *tail_bwd_ms = c->t_tail_bwd;
  return 0;
}

int lapgpu_last_timings(lapgpu_ctx* c, double* spmv_ms, double* ldl_ms,
                        double* vec_ms, double* total_ms, int64_t* its) {
  if (!c) return -1;
  if (spmv_ms) *spmv_ms = c->t_spmv;
  if (ldl_ms) *ldl_ms = c->t_ldl;
  if (vec_ms) *vec_ms = c->t_vec;
  if (total_ms) *total_ms = c->t_total;
  if (its) *its = c->t_its;
  return 0;
}

}  // extern "C"
