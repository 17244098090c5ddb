// lapgpu_device.hpp — device helpers shared by the gfx950 kernels: constants,
// the solve state, wave/block reductions and the slot-program runner.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lapgpu {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kRedBlocks = 1024;  // fixed grid of the reduction kernels
#ifndef LAPGPU_TAIL_THREADS
#define LAPGPU_TAIL_THREADS 512
#endif
constexpr int kTailThreads = LAPGPU_TAIL_THREADS;  // = kTailLanes of the plan
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

// MI355X deals workgroups to its 8 XCDs round-robin (workgroup b to XCD
// b mod 8), and each XCD has its own L2.  A kernel that gathers takes its
// rows through xcd_block(): the logical block of this workgroup, chosen so
// that each XCD runs one contiguous range of logical blocks, so that the
// rows an XCD holds at once are neighbours and share their gathered lines
// in its L2 instead of every XCD fetching them.  A bijection of [0,
// gridDim.x): logical block l computes what workgroup l computed before
// (same rows, same partial-sum slot), so results do not change.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, nb = gridDim.x;
  const int x = b % kXcds, per = nb / kXcds, rem = nb % kXcds;
  return x * per + min(x, rem) + b / kXcds;
}

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
  double bmu;      // mean of b removed by the closure (0 without centering)
  double true_nr;  // after the solve: ||b - bmu - la x|| / nb in FP64
};

// ---------------------------------------------------------------------------
// reductions

template <int G, typename V>
__device__ __forceinline__ V group_sum(V v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

// sum over k = k0 + lane, k0 + lane + G, ... < k1 of w[k] * x(idx[k]), added
// in that order.  The index and weight loads of U consecutive terms are
// issued together, then their U gathers, so a lane waits about two memory
// latencies per U terms instead of two per term (the sums are the same as
// the one-term-at-a-time loop, bit for bit).
template <int G, int U, typename T, typename X>
__device__ __forceinline__ T strided_dot(int k0, int k1, int lane, const int* __restrict__ idx,
                                         const T* __restrict__ w, X x) {
  T acc = 0;
  for (int kb = k0 + lane; kb < k1; kb += U * G) {
    int ii[U];
    T ww[U], xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kb + u * G;
      if (k < k1) {
        ii[u] = idx[k];
        ww[u] = w[k];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (kb + u * G < k1) xv[u] = x(ii[u]);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (kb + u * G < k1) acc += ww[u] * xv[u];
  }
  return acc;
}

// A grid-stride loop over rows (this lane group's rows first_row,
// first_row + stride, ...), G lanes per row, with RU rows in flight: the
// row pointers of RU rows are loaded together, then their first terms
// (entry k0 + lane), then those terms' gathers, so a row costs about three
// memory latencies per RU rows instead of three per row.  Each row's sum is
// the same as the one-row-at-a-time loop's (terms k0 + lane, k0 + lane + G,
// ... added in that order, then group_sum); emit(row, sum) is called for
// the rows in order.
template <int G, int RU, typename T, typename X, typename Emit>
__device__ __forceinline__ void rows_in_flight(int first_row, int stride, int n,
                                               const int* __restrict__ ptr,
                                               const int* __restrict__ idx,
                                               const T* __restrict__ w, X x, Emit emit) {
  const int lane = threadIdx.x % G;
  for (int r0 = first_row; r0 < n; r0 += RU * stride) {
    int k0[RU], k1[RU], ii[RU];
    T ww[RU], xv[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int row = r0 + u * stride;
      k0[u] = row < n ? ptr[row] : 0;
      k1[u] = row < n ? ptr[row + 1] : 0;
    }
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int k = k0[u] + lane;
      if (k < k1[u]) {
        ii[u] = idx[k];
        ww[u] = w[k];
      }
    }
#pragma unroll
    for (int u = 0; u < RU; ++u)
      if (k0[u] + lane < k1[u]) xv[u] = x(ii[u]);
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int row = r0 + u * stride;
      if (row >= n) break;  // uniform over the lane group
      T acc = 0;
      if (k0[u] + lane < k1[u]) acc += ww[u] * xv[u];
      for (int k = k0[u] + lane + G; k < k1[u]; k += G) acc += w[k] * x(idx[k]);
      acc = group_sum<G>(acc);
      emit(row, acc);
    }
  }
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
// Slot programs.  The narrow tail of the DAG (the last levels, at most
// kTailCapV vertices) is run by one workgroup that holds y of all tail
// vertices in LDS and does the forward sweep, the diagonal scaling and the
// backward sweep of those levels with one workgroup barrier per level; the
// batched solver runs whole graphs the same way.  The work of each level is
// a host-packed "slot program" (TailSlot / TailStep of lapgpu_plan.hpp):
// lane t of step j sums E = 2^elog products w * y[src] (the step's planes),
// the lanes of an aligned power-of-two segment are summed with DPP row
// shifts so that the segment's LAST lane holds the sum, and that lane adds
// it to y[row].  Several entries per lane cut the cross-lane stages, which
// are the VALU cost of a step, and the lanes a level needs.
//
// Latency: every slot load is unconditional (inactive lanes re-read a
// valid slot of the same step), so the compiler can wait on exactly the
// loads it needs; slots are loaded D steps ahead, and the step table 64
// steps ahead with one vector load per lane, read out with readlane (no
// scalar loads in the loop, whose counter is shared with LDS).  A step's
// largest segment bounds the DPP stages run.

constexpr int kTailCapV = 18432;
// vertices of the single-workgroup tail of a context (= the planner's kTailCap)
#ifndef LAPGPU_TAIL_CAP
#define LAPGPU_TAIL_CAP 18432
#endif
constexpr int kTailCapLds = LAPGPU_TAIL_CAP;
// partial slots of a slot program, past its own indices (= the planner's
// kPartCap): the y arrays of the programs hold kTailCapLds + kPartSlots
constexpr int kPartSlots = 1024;

struct alignas(16) TailSlotD {
  double w;
  unsigned src, meta;  // meta: row (26 bits) | seglog << 26 | lead << 29
};
struct alignas(4) TailSlotF {
  float w;
  unsigned src, meta;
};

// off = first slot; meta = count | end << 16 | elog << 17 | maxlog << 19
struct alignas(8) TailStepD {
  int off, meta;
};

__device__ __forceinline__ int2 load_step(const TailStepD* __restrict__ steps, int nsteps, int j) {
  const int jj = j < nsteps ? j : (nsteps > 0 ? nsteps - 1 : 0);
  const int2 v = *reinterpret_cast<const int2*>(steps + jj);
  return j < nsteps ? v : make_int2(0, 0);
}

// lane t's EM planes of the step (off, meta), addressed from a wave-uniform
// base per plane; lanes >= count re-read a valid slot of the same step
template <int EM, typename S>
__device__ __forceinline__ void load_planes(S (&r)[EM], const S* __restrict__ slots, int off,
                                            int meta, int t) {
  const int cnt = meta & 0xffff;
  const unsigned tt = min((unsigned)t, (unsigned)max(cnt - 1, 0));
#pragma unroll
  for (int e = 0; e < EM; ++e) {
    const S* __restrict__ base = slots + (off + e * cnt);
    r[e] = base[tt];
  }
}

template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xF, true);
  return __hiloint2double(hi, lo);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, true));
}

// Sum over aligned segments of 2^seglog lanes; the last lane of each segment
// receives the segment's sum (other lanes hold partial sums).  Row shifts
// 1, 2, 4, 8 inside each 16-lane row, then row_bcast:15 (lane 15 of a row
// into the next row, rows 1 and 3) and row_bcast:31 (lane 31 into rows 2
// and 3); lanes a DPP move does not write read 0.  Stages at or above the
// wave-uniform bound maxlog are skipped.
template <typename Y>
__device__ __forceinline__ Y seg_sum_last(Y v, int seglog, int maxlog) {
  Y t;
  if (maxlog >= 1) {
    t = dpp_mov<0x111, 0xF>(v);  // row_shr:1
    if (seglog >= 1) v += t;
  }
  if (maxlog >= 2) {
    t = dpp_mov<0x112, 0xF>(v);  // row_shr:2
    if (seglog >= 2) v += t;
  }
  if (maxlog >= 3) {
    t = dpp_mov<0x114, 0xF>(v);  // row_shr:4
    if (seglog >= 3) v += t;
  }
  if (maxlog >= 4) {
    t = dpp_mov<0x118, 0xF>(v);  // row_shr:8
    if (seglog >= 4) v += t;
  }
  if (maxlog >= 5) {
    t = dpp_mov<0x142, 0xA>(v);  // row_bcast:15
    if (seglog >= 5) v += t;
  }
  if (maxlog >= 6) {
    t = dpp_mov<0x143, 0xC>(v);  // row_bcast:31
    if (seglog >= 6) v += t;
  }
  return v;
}

// one step for lane threadIdx.x: sum of its E products, segment sum, lead add
template <bool LDSY, int EM, typename S, typename Y>
__device__ __forceinline__ void run_step(const S (&r)[EM], int meta, Y* ly) {
  const int count = meta & 0xffff;
  if ((int)threadIdx.x < count) {  // count is a multiple of 64: wave-uniform
    Y v = r[0].w * ly[r[0].src];
#pragma unroll
    for (int e = 1; e < EM; ++e) v += r[e].w * ly[r[e].src];
    v = seg_sum_last(v, (int)((r[0].meta >> 26) & 7u), (meta >> 19) & 7);
    // in LDS the lead lanes add without waiting for a reply: a single-
    // segment row has exactly one add per level (the same result as a
    // read-modify-write), the chunks of a split row either store partial
    // sums (lead 3, summed in order by a later fix-up step) or add in any
    // order (lead 2).  In global memory a single-segment row is read-
    // modified-written, so the workgroup's later loads see it through its L1.
    const unsigned lead = r[0].meta >> 29;
    const unsigned row = r[0].meta & ((1u << 26) - 1);
    if (lead == 3) {
      ly[row] = v;
    } else if (LDSY) {
      if (lead != 0) atomicAdd(&ly[row], v);
    } else if (lead == 1) {
      ly[row] += v;
    } else if (lead == 2) {
      atomicAdd(&ly[row], v);
    }
  }
}

// Workgroup barrier for LDS-only communication: waits on LDS operations,
// not on global loads in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Runs nsteps steps of a slot program on ly (in LDS when LDSY, else in
// global memory).  Every step of the program has E = EM entries per lane.
// Slots are double-buffered D steps at a time in two register sets used in
// turn (no copies between them): the loads of the next D steps are issued
// before the current D steps run, so every wait counts exactly the loads
// still needed.  2D divides 64, so a block of the step table is consumed
// by whole iterations.
template <int D, bool LDSY, int EM, typename S, typename Y>
__device__ __forceinline__ void run_program(const TailStepD* __restrict__ steps, int nsteps,
                                            const S* __restrict__ slots, Y* ly) {
  static_assert(64 % (2 * D) == 0, "2D divides 64");
  const int t = threadIdx.x;
  const int lane = t & 63;
  // step table: blkA = steps [base, base + 64), blkB = the next 64
  int2 blkA = load_step(steps, nsteps, lane);
  int2 blkB = load_step(steps, nsteps, 64 + lane);
  S a[D][EM], b[D][EM];
#pragma unroll
  for (int d = 0; d < D; ++d)
    load_planes(a[d], slots, __builtin_amdgcn_readlane(blkA.x, d),
                __builtin_amdgcn_readlane(blkA.y, d), t);
  for (int c = 0; c < nsteps; c += 2 * D) {
    const int i0 = c & 63;  // steps c .. c+2D-1 are blkA[i0 .. i0+2D)
    // steps c+2D .. c+3D-1: blkA, or blkB when c+2D starts a new block
    const bool nb = i0 + 2 * D == 64;
#pragma unroll
    for (int d = 0; d < D; ++d)
      load_planes(b[d], slots, __builtin_amdgcn_readlane(blkA.x, i0 + D + d),
                  __builtin_amdgcn_readlane(blkA.y, i0 + D + d), t);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int meta = __builtin_amdgcn_readlane(blkA.y, i0 + d);
      run_step<LDSY>(a[d], meta, ly);
      if ((meta >> 16) & 1) {
        if (LDSY)
          lds_barrier();
        else
          __syncthreads();
      }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int idx = (i0 + 2 * D + d) & 63;
      const int off = nb ? __builtin_amdgcn_readlane(blkB.x, idx) : __builtin_amdgcn_readlane(blkA.x, idx);
      const int meta = nb ? __builtin_amdgcn_readlane(blkB.y, idx) : __builtin_amdgcn_readlane(blkA.y, idx);
      load_planes(a[d], slots, off, meta, t);
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int meta = __builtin_amdgcn_readlane(blkA.y, i0 + D + d);
      run_step<LDSY>(b[d], meta, ly);
      if ((meta >> 16) & 1) {
        if (LDSY)
          lds_barrier();
        else
          __syncthreads();
      }
    }
    if (nb) {
      blkA = blkB;
      blkB = load_step(steps, nsteps, c + 2 * D + 64 + lane);
    }
  }
}

// Steps per register set for E entries per lane (2 D E slots per lane in
// registers).  Measured on MI355X with 512 tail threads and E = 4 (grid3(128)
// tail per iteration; path + 4 matchings n = 1e6 FP64 tail): depth 4 (D = 1)
// 0.96 / 5.55 ms, 8 (D = 2) 0.94 / 5.73 ms, 16 (D = 4) 1.10 / 6.64 ms; the
// tail's steps are not bound by their slot loads.
#ifndef LAPGPU_TAIL_DEPTH
#define LAPGPU_TAIL_DEPTH 8
#endif
template <int E>
__host__ __device__ constexpr int tail_depth() {
  return (kTailThreads <= 256 ? 16 : LAPGPU_TAIL_DEPTH) / E;
}

}  // namespace lapgpu
