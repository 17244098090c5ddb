// lapgpu_chain.hpp — runs of narrow level groups in one launch.
//
// Near the tail the level groups are narrow (tens to a thousand rows): such
// a launch does not fill the device, and its time is the launch itself
// plus one row's chain of dependent loads (row pointer, indices, gathers,
// store).  In the single right-hand-side sweeps a run of consecutive
// narrow groups is therefore run by ONE launch.  Its work is cut into tickets (a range of a
// group's light rows, or its heavy rows), in sweep order, so every row a
// ticket reads was computed by an earlier ticket or before the launch.
// Workgroups take tickets from an agent-scope counter, so tickets start in
// order whatever the dispatch order, and the lowest unfinished ticket always
// belongs to a running workgroup whose inputs are all done: the launch
// cannot deadlock, also beside other kernels.  A ticket's index and weight
// loads do not depend on earlier groups, so they overlap the wait.
//
// Readiness is carried by the values themselves: before the sweeps the
// positions the runs write hold a sentinel (a signalling NaN, which
// arithmetic never produces); a row's result is written with an
// agent-scope store (written through, visible to every XCD), and a gather
// that reads the sentinel polls the position with agent-scope loads until
// the value arrives.  A value is written once per sweep, so a gather sees
// either the sentinel or the final value.  Rows are summed exactly as the
// one-launch-per-group kernels sum them (same lanes per row, term order and
// reduction tree): the results are the same bits.  Every wait is bounded;
// one that runs out sets ctl[0] and stops all waits, and the host reports
// the call as failed and goes back to one launch per group.
//
// Wide groups keep a launch of their own: there the rows of a group depend
// on rows spread over the whole previous group, so tickets taken ahead of
// it would mostly wait, holding CU slots (measured: every group of a sweep
// in one such launch was slower than one launch per group on every graph).
// So do the multi-RHS sweeps, whose K-column rows are a wave each: chained
// runs of their narrow groups measured slower at every threshold.
#pragma once

#include "lapgpu_kernels.hpp"

namespace lapgpu {

// One ticket: light rows [lo, hi) of a group with G lanes per row (rows
// longer than heavy_limit<G> skipped), or, with G = 0, heavy rows
// heavy[lo..hi) (up to four, one wave each).  Shared by host and device.
struct ChainTicket {
  int32_t lo, hi, G, pad;
};

// Ticket sizes: a group's light rows are cut into ranges of whole passes of
// the workgroup, as many passes as keep about kChainTickets tickets per
// group (at most kChainPasses); a run's launch has at most kChainGrid
// workgroups.
constexpr int kChainTickets = 64;
constexpr int kChainPasses = 8;
constexpr int kChainGrid = 512;

// ctl[0]: a wait ran out (sticky until the host clears it); ctl[1]: the
// polls one wait may take (set by the host: kChainPollMax, or fewer to
// exercise the fallback); ctl[2 + r]: ticket counter of run r (set to
// zero before the sweeps)
constexpr int kChainPollMax = 1 << 18;  // ~0.3 s

constexpr unsigned long long kSentD = 0x7ff4dead5e47f00dULL;
constexpr unsigned int kSentF = 0x7fa5e47fu;

template <typename T>
struct Sent;
template <>
struct Sent<double> {
  __device__ static bool is(double v) {
    return (unsigned long long)__double_as_longlong(v) == kSentD;
  }
  __device__ static double value() { return __longlong_as_double((long long)kSentD); }
  __device__ static double load(const double* p) {
    return __longlong_as_double(
        __hip_atomic_load((long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  __device__ static void store(double* p, double v) {
    if (is(v)) v = __longlong_as_double(0x7ff8000000000000LL);
    __hip_atomic_store((long long*)p, __double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
};
template <>
struct Sent<float> {
  __device__ static bool is(float v) { return (unsigned int)__float_as_int(v) == kSentF; }
  __device__ static float value() { return __int_as_float((int)kSentF); }
  __device__ static float load(const float* p) {
    return __int_as_float(
        __hip_atomic_load((int*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  __device__ static void store(float* p, float v) {
    if (is(v)) v = __int_as_float(0x7fc00000);
    __hip_atomic_store((int*)p, __float_as_int(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
};

#ifndef LAPGPU_CHAIN_SLEEP
#define LAPGPU_CHAIN_SLEEP 32  // units of 64 clocks between polls
#endif

// Waits until none of this lane's values xv[u] with bit u of `pend` set
// (read from addr[u]) is the sentinel any more: the pending values are
// polled together until they arrive.  A wait that runs out (or finds
// another wait of the launch ran out) sets ctl[0] and leaves zeros.
template <int U, typename T>
__device__ __forceinline__ void chain_poll(T (&xv)[U], const T* const (&addr)[U], unsigned pend,
                                           int* ctl) {
  const int pmax = ctl[1];
  for (int polls = 1; pend; ++polls) {
    __builtin_amdgcn_s_sleep(LAPGPU_CHAIN_SLEEP);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (pend & (1u << u)) xv[u] = Sent<T>::load(addr[u]);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if ((pend & (1u << u)) && !Sent<T>::is(xv[u])) pend &= ~(1u << u);
    if (pend && (polls >= pmax ||
                 ((polls & 63) == 0 &&
                  __hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0))) {
      __hip_atomic_store(ctl, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (pend & (1u << u)) xv[u] = (T)0;
      pend = 0;
    }
  }
}

// The ticket loop shared by the chained kernels: tickets of this run in
// order, the next one claimed while the current one runs.
template <typename F>
__device__ __forceinline__ void chain_tickets(const ChainTicket* __restrict__ tickets, int t0,
                                              int t1, int* ctr, F&& run) {
  __shared__ int next[2];
  if (threadIdx.x == 0)
    next[0] = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  int t = next[0];
  int par = 1;
  while (t0 + t < t1) {
    int claim = 0;
    if (threadIdx.x == 0)
      claim = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    run(tickets[t0 + t]);
    if (threadIdx.x == 0) next[par] = claim;
    __syncthreads();
    t = next[par];
    par ^= 1;
  }
}

// ---- single right-hand side (k_group's rows) ---------------------------------

// strided_dot<G, 8> over a composite row whose A terms may still be pending:
// the same loads, then the wait for pending ones, then the same sums
template <int G, typename T, typename TB>
__device__ __forceinline__ T chain_dot(int k0, int k1, int lane, int n,
                                       const int* __restrict__ idx, const T* __restrict__ w,
                                       const T* A, const TB* B, int* ctl) {
  constexpr int U = 8;
  T acc = 0;
  for (int kb = k0 + lane; kb < k1; kb += U * G) {
    int ii[U];
    T ww[U], xv[U];
    const T* addr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kb + u * G;
      if (k < k1) {
        ii[u] = idx[k];
        ww[u] = w[k];
      }
    }
    unsigned pend = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      addr[u] = A;
      if (kb + u * G < k1) {
        if (ii[u] < n) {
          addr[u] = A + ii[u];
          xv[u] = *addr[u];
        } else {
          xv[u] = (T)B[ii[u] - n];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (kb + u * G < k1 && ii[u] < n && Sent<T>::is(xv[u])) pend |= 1u << u;
    if (pend) chain_poll<U>(xv, addr, pend, ctl);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (kb + u * G < k1) acc += ww[u] * xv[u];
  }
  return acc;
}

template <int G, typename T, typename TB>
__device__ __forceinline__ void chain_light(const ChainTicket& tk, int n,
                                            const int* __restrict__ ptr,
                                            const int* __restrict__ idx,
                                            const T* __restrict__ w, const T* A, const TB* B,
                                            T* out, int* ctl) {
  const int lane = threadIdx.x % G;
  for (int row = tk.lo + (int)threadIdx.x / G; row < tk.hi; row += kBlock / G) {
    const int k0 = ptr[row], k1 = ptr[row + 1];
    if (k1 - k0 > heavy_limit<G>()) continue;  // the same for the G lanes of the row
    T acc = chain_dot<G>(k0, k1, lane, n, idx, w, A, B, ctl);
    acc = group_sum<G>(acc);
    if (lane == 0) Sent<T>::store(out + row, acc);
  }
}

template <typename T, typename TB>
__device__ __forceinline__ void chain_heavy(const ChainTicket& tk, int n,
                                            const int* __restrict__ heavy,
                                            const int* __restrict__ ptr,
                                            const int* __restrict__ idx,
                                            const T* __restrict__ w, const T* A, const TB* B,
                                            T* out, int* ctl) {
  const int h = tk.lo + (int)threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (h >= tk.hi) return;
  const int row = heavy[h];
  T acc = chain_dot<kWave>(ptr[row], ptr[row + 1], lane, n, idx, w, A, B, ctl);
  acc = group_sum<kWave>(acc);
  if (lane == 0) Sent<T>::store(out + row, acc);
}

// Tickets [t0, t1) of one run (counter ctl[2 + run]):
//   out[row] = sum_k w[k] * (idx[k] < n ? A[idx[k]] : B[idx[k] - n])  as k_group
template <typename T, typename TB>
__global__ void __launch_bounds__(kBlock)
k_chain(int n, const ChainTicket* __restrict__ tickets, int t0, int t1, int* ctl, int run,
        const int* __restrict__ heavy, const int* __restrict__ ptr,
        const int* __restrict__ idx, const T* __restrict__ w, const T* A, const TB* B,
        T* out) {
  chain_tickets(tickets, t0, t1, ctl + 2 + run, [&](const ChainTicket& tk) {
    switch (tk.G) {
      case 4: chain_light<4>(tk, n, ptr, idx, w, A, B, out, ctl); break;
      case 8: chain_light<8>(tk, n, ptr, idx, w, A, B, out, ctl); break;
      case 16: chain_light<16>(tk, n, ptr, idx, w, A, B, out, ctl); break;
      case 32: chain_light<32>(tk, n, ptr, idx, w, A, B, out, ctl); break;
      default: chain_heavy(tk, n, heavy, ptr, idx, w, A, B, out, ctl); break;
    }
  });
}

// Before a pair of sweeps: the sentinel into both sweeps' outputs over
// elements [lo, hi) (the positions the runs write; K interleaved columns:
// lo K .. hi K) and the runs' ticket counters to zero.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_chain_prep(int64_t lo, int64_t hi, T* y, T* z, int* ctl, int nctr) {
  const T s = Sent<T>::value();
  for (int64_t i = lo + (int64_t)blockIdx.x * kBlock + threadIdx.x; i < hi;
       i += (int64_t)gridDim.x * kBlock) {
    y[i] = s;
    z[i] = s;
  }
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < nctr; i += kBlock) ctl[2 + i] = 0;
}

}  // namespace lapgpu
