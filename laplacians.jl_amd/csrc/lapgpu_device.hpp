// lapgpu_device.hpp — device helpers shared by the gfx950 kernels: constants,
// the solve state, wave/block reductions and the slot-program runner.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lapgpu {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kRedBlocks = 1024;  // fixed grid of the reduction kernels
constexpr int kTailThreads = 1024;
constexpr double kEps = 2.220446049250313e-16;  // eps(Float64)

// partial-sum slots (each kRedBlocks doubles)
enum Slot {
  S_SUM = 0,   // sum of a vector (means)
  S_RR = 1,    // r.r
  S_PQ = 2,    // p.q
  S_ZR = 3,    // z.r
  S_AP = 4,    // sum|p|  (al check, before x update)
  S_AX = 5,    // sum|x|
  S_AQ = 6,    // sum|q|
  S_AR = 7,    // sum|r|  (before r update)
  S_AZ = 8,    // sum|z|
  S_AP2 = 9,   // sum|p|  (beta check)
  kNumSlots = 10
};

struct DevState {
  double nb, rho, best_rho, al, beta, nr, bestnr, mu;
  double tol, stag_fac, rho0, bestnr_prev, switch_tol;
  int64_t itcnt, maxits, stag_test, stag_count;
  uint64_t t1, maxticks;
  int32_t stopped, stop_reason, pending_copy, check10;
  int32_t center, has_maxtime;
  int32_t fp32;    // FP32-inner mode: convergence needs an FP64 true residual
  int32_t verify;  // set when the recursive residual passed tol (FP32 mode)
  int32_t spmv64;  // FP32 mode: the SpMV has switched to FP64
  int32_t pad2;
};

// ---------------------------------------------------------------------------
// reductions

template <int G, typename V>
__device__ __forceinline__ V group_sum(V v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

// Sum NV values over the block; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double red[NV][kBlock / kWave];
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = group_sum<kWave>(v[i]);
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) red[i][wid] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int w = 0; w < kBlock / kWave; ++w) s += red[i][w];
      v[i] = s;
    }
  }
}

// Sum of the kRedBlocks partials of one slot by a block of kScalarThreads
// (= kRedBlocks) threads, one partial each, in a fixed tree order.  The
// result is returned to every thread.
constexpr int kScalarThreads = kRedBlocks;
__device__ __forceinline__ double wave_total(const double* part, int slot) {
  __shared__ double red[kScalarThreads / kWave];
  double v = part[slot * kRedBlocks + threadIdx.x];
  v = group_sum<kWave>(v);
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kScalarThreads / kWave; ++w) s += red[w];
  return s;
}

// ---------------------------------------------------------------------------
// The narrow tail of the DAG (the last levels, at most kTailCapV vertices):
// one workgroup holds y of all tail vertices in LDS and runs the forward
// sweep, the diagonal scaling and the backward sweep of those levels with a
// workgroup barrier per level.  The work of each level is a host-packed
// "slot program" (TailSlot / TailStep of lapgpu_plan.hpp): lane t of step j
// computes w * y[src], the lanes of an aligned power-of-two segment are
// summed with xor-shuffles, and the segment's lead lane adds the sum to
// y[row].  Slots are loaded kTailDepth steps ahead of their use.

constexpr int kTailCapV = 18432;
constexpr int kTailDepth = 8;

struct alignas(16) TailSlotD {
  double w;
  unsigned src, meta;  // meta: row (26 bits) | seglog << 26 | lead << 29
};
struct alignas(4) TailSlotF {
  float w;
  unsigned src, meta;
};

struct alignas(16) TailStepD {
  int off, count, end, pad;
};

template <typename S>
__device__ __forceinline__ S load_slot(const S* __restrict__ slots, const TailStepD& st, int t) {
  S z;
  if (t < st.count) {
    z = slots[st.off + t];
  } else {
    z.w = 0;
    z.src = 0;
    z.meta = 0;
  }
  return z;
}

template <typename S, typename Y>
__device__ __forceinline__ void run_slot(const S& sl, const TailStepD& st, Y* ly) {
  if ((int)threadIdx.x < st.count) {
    Y v = sl.w * ly[sl.src];
    const int seg = 1 << ((sl.meta >> 26) & 7u);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const Y t = __shfl_xor(v, o, kWave);
      if (o < seg) v += t;
    }
    const unsigned lead = sl.meta >> 29;
    const unsigned row = sl.meta & ((1u << 26) - 1);
    if (lead == 1)
      ly[row] += v;
    else if (lead == 2)
      atomicAdd(&ly[row], v);
  }
}

// Runs nsteps steps of a slot program on ly, loading slots D steps ahead.
template <int D = kTailDepth, typename S, typename Y>
__device__ __forceinline__ void run_program(const TailStepD* __restrict__ steps,
                                            int nsteps, const S* __restrict__ slots,
                                            Y* ly) {
  constexpr int kTailDepth = D;
  const int t = threadIdx.x;
  TailStepD ia[kTailDepth], ib[kTailDepth];
  S a[kTailDepth], b[kTailDepth];
  const TailStepD none = {0, 0, 0, 0};
#pragma unroll
  for (int d = 0; d < kTailDepth; ++d) {
    ia[d] = d < nsteps ? steps[d] : none;
    a[d] = load_slot(slots, ia[d], t);
  }
  for (int c = 0; c < nsteps; c += kTailDepth) {
#pragma unroll
    for (int d = 0; d < kTailDepth; ++d) {
      const int j = c + kTailDepth + d;
      ib[d] = j < nsteps ? steps[j] : none;
      b[d] = load_slot(slots, ib[d], t);
    }
#pragma unroll
    for (int d = 0; d < kTailDepth; ++d) {
      if (c + d < nsteps) {
        run_slot(a[d], ia[d], ly);
        if (ia[d].end) __syncthreads();
      }
    }
#pragma unroll
    for (int d = 0; d < kTailDepth; ++d) {
      a[d] = b[d];
      ia[d] = ib[d];
    }
  }
}

}  // namespace lapgpu
