// lapgpu_kernels.hpp — HIP kernels for gfx950 (CDNA4, wave64) of the
// approxchol_lap solve phase.  Included by lapgpu.hip only.
//
// Layout: all vectors are in "level order" (see lapgpu_plan.cpp): position
// pos holds vertex perm[pos].  The reductions are deterministic: every
// reduction kernel runs a fixed grid of kRedBlocks blocks and writes one
// partial per block; a single-wave scalar kernel adds the partials in a fixed
// tree order.  All loop kernels return at once when the solve has stopped,
// so the host can enqueue iterations ahead of the stopping decision.
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
  double tol, stag_fac, rho0;
  int64_t itcnt, maxits, stag_test, stag_count;
  uint64_t t1, maxticks;
  int32_t stopped, stop_reason, pending_copy, check10;
  int32_t center, has_maxtime, pad0, pad1;
};

// ---------------------------------------------------------------------------
// reductions

template <int G>
__device__ __forceinline__ double group_sum(double v) {
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
// permutation and setup

__global__ void k_gather_perm(int n, const int* __restrict__ perm,
                              const double* __restrict__ src,
                              double* __restrict__ dst,
                              double* __restrict__ part) {
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double t = src[perm[i]];
    dst[i] = t;
    v[0] += t;
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) part[S_SUM * kRedBlocks + blockIdx.x] = v[0];
}

__global__ void k_scatter_perm(int n, const int* __restrict__ perm,
                               const double* __restrict__ src,
                               double* __restrict__ dst) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    dst[perm[i]] = src[i];
}

// mean of the gathered rhs
__global__ void k_scalar_mean(int n, const double* __restrict__ part,
                              DevState* st) {
  const double s = wave_total(part, S_SUM);
  if (threadIdx.x == 0) st->mu = st->center ? s / (double)n : 0.0;
}

// r = b - mean(b); x = bestx = 0; partial r.r
__global__ void k_center(int n, const double* __restrict__ b,
                         double* __restrict__ r, double* __restrict__ x,
                         double* __restrict__ bestx, double* __restrict__ part,
                         const DevState* __restrict__ st) {
  const double mu = st->mu;
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double t = b[i] - mu;
    r[i] = t;
    x[i] = 0.0;
    bestx[i] = 0.0;
    v[0] += t * t;
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) part[S_RR * kRedBlocks + blockIdx.x] = v[0];
}

__global__ void k_scalar_nb(const double* __restrict__ part, DevState* st) {
  const double rr = wave_total(part, S_RR);
  if (threadIdx.x == 0) {
    st->nb = sqrt(rr);
    if (st->nb == 0.0) {
      st->stopped = 1;
      st->stop_reason = 9;  // zero rhs
    }
  }
}

// ---------------------------------------------------------------------------
// LDLinv sweeps (gathers over one level; G lanes per row)

template <int G>
__global__ void __launch_bounds__(kBlock)
k_fwd_level(int lo, int hi, const int* __restrict__ Fptr,
            const int* __restrict__ Fidx, const double* __restrict__ Fw,
            const double* __restrict__ r, double* y,
            const DevState* __restrict__ st) {
  const int gid = blockIdx.x * kBlock + threadIdx.x;
  const int row = lo + gid / G;
  const int lane = gid % G;
  if (row >= hi) return;
  const int k0 = Fptr[row], k1 = Fptr[row + 1];
  double acc = 0.0;
  for (int k = k0 + lane; k < k1; k += G) acc += Fw[k] * y[Fidx[k]];
  acc = group_sum<G>(acc);
  if (lane == 0 && !st->stopped) y[row] = r[row] + acc;
}

template <int G>
__global__ void __launch_bounds__(kBlock)
k_bwd_level(int lo, int hi, const int* __restrict__ Bptr,
            const int* __restrict__ Bidx, const double* __restrict__ Bw,
            const double* __restrict__ s, double* y,
            const DevState* __restrict__ st) {
  const int gid = blockIdx.x * kBlock + threadIdx.x;
  const int row = lo + gid / G;
  const int lane = gid % G;
  if (row >= hi) return;
  const int k0 = Bptr[row], k1 = Bptr[row + 1];
  double acc = 0.0;
  for (int k = k0 + lane; k < k1; k += G) acc += Bw[k] * y[Bidx[k]];
  acc = group_sum<G>(acc);
  if (lane == 0 && !st->stopped) y[row] = s[row] * y[row] + acc;
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
  unsigned short src, row;
  unsigned char seglog, lead;
  unsigned short pad;
};
struct alignas(16) TailStepD {
  int off, count, end, pad;
};

__device__ __forceinline__ TailSlotD load_slot(const TailSlotD* __restrict__ slots,
                                               const TailStepD& st, int t) {
  TailSlotD z;
  if (t < st.count) {
    z = slots[st.off + t];
  } else {
    z.w = 0.0;
    z.src = 0;
    z.row = 0;
    z.seglog = 0;
    z.lead = 0;
    z.pad = 0;
  }
  return z;
}

__device__ __forceinline__ void run_slot(const TailSlotD& sl, const TailStepD& st,
                                         double* ly) {
  if ((int)threadIdx.x < st.count) {
    double v = sl.w * ly[sl.src];
    const int seg = 1 << sl.seglog;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const double t = __shfl_xor(v, o, kWave);
      if (o < seg) v += t;
    }
    if (sl.lead == 1)
      ly[sl.row] += v;
    else if (sl.lead == 2)
      atomicAdd(&ly[sl.row], v);
  }
}

// Runs nsteps steps of a slot program on ly.
__device__ __forceinline__ void run_program(const TailStepD* __restrict__ steps,
                                            int nsteps,
                                            const TailSlotD* __restrict__ slots,
                                            double* ly) {
  const int t = threadIdx.x;
  TailStepD ia[kTailDepth], ib[kTailDepth];
  TailSlotD a[kTailDepth], b[kTailDepth];
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

// ext[i] = r[T0+i] + sum of the forward in-edges of tail row i that come
// from outside the tail (G lanes per row)
template <int G>
__global__ void __launch_bounds__(kBlock)
k_tail_ext(int T0, int NT, const int* __restrict__ Eptr,
           const int* __restrict__ Eidx, const double* __restrict__ Ew,
           const double* __restrict__ r, const double* __restrict__ y,
           double* __restrict__ ext, const DevState* __restrict__ st) {
  const int gid = blockIdx.x * kBlock + threadIdx.x;
  const int row = gid / G;
  const int lane = gid % G;
  if (row >= NT) return;
  const int k0 = Eptr[row], k1 = Eptr[row + 1];
  double acc = 0.0;
  for (int k = k0 + lane; k < k1; k += G) acc += Ew[k] * y[Eidx[k]];
  acc = group_sum<G>(acc);
  if (lane == 0 && !st->stopped) ext[row] = r[T0 + row] + acc;
}

// forward + diagonal + backward on the tail; writes yF of the tail to y
__global__ void __launch_bounds__(kTailThreads)
k_tail(int T0, int NT, const double* __restrict__ ext,
       const double* __restrict__ s, double* __restrict__ y,
       const TailStepD* __restrict__ fsteps, int nfsteps,
       const TailSlotD* __restrict__ fslots,
       const TailStepD* __restrict__ bsteps, int nbsteps,
       const TailSlotD* __restrict__ bslots,
       const DevState* __restrict__ st) {
  __shared__ double ly[kTailCapV];
  if (st->stopped) return;
  for (int i = threadIdx.x; i < NT; i += kTailThreads) ly[i] = ext[i];
  __syncthreads();
  run_program(fsteps, nfsteps, fslots, ly);
  __syncthreads();
  for (int i = threadIdx.x; i < NT; i += kTailThreads) ly[i] *= s[T0 + i];
  __syncthreads();
  run_program(bsteps, nbsteps, bslots, ly);
  __syncthreads();
  for (int i = threadIdx.x; i < NT; i += kTailThreads) y[T0 + i] = ly[i];
}

// partial sum of y (for the mean removal of LDLsolver)
__global__ void k_vsum(int n, const double* __restrict__ y,
                       double* __restrict__ part,
                       const DevState* __restrict__ st) {
  if (st->stopped) return;
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    v[0] += y[i];
  block_sum<1>(v);
  if (threadIdx.x == 0) part[S_SUM * kRedBlocks + blockIdx.x] = v[0];
}

__global__ void k_scalar_mu(int n, const double* __restrict__ part,
                            DevState* st) {
  if (st->stopped) return;
  const double s = wave_total(part, S_SUM);
  if (threadIdx.x == 0) st->mu = s / (double)n;
}

// z = y - mu; partials z.r, and (every 10th iteration) sum|z|, sum|p|
__global__ void k_zdot(int n, const double* __restrict__ y,
                       const double* __restrict__ r,
                       const double* __restrict__ p, double* __restrict__ z,
                       double* __restrict__ part,
                       const DevState* __restrict__ st) {
  if (st->stopped) return;
  const double mu = st->mu;
  const bool chk = st->check10 != 0;
  double v[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double t = y[i] - mu;
    z[i] = t;
    v[0] += t * r[i];
    if (chk) {
      v[1] += fabs(t);
      v[2] += fabs(p[i]);
    }
  }
  block_sum<3>(v);
  if (threadIdx.x == 0) {
    part[S_ZR * kRedBlocks + blockIdx.x] = v[0];
    part[S_AZ * kRedBlocks + blockIdx.x] = v[1];
    part[S_AP2 * kRedBlocks + blockIdx.x] = v[2];
  }
}

// rho of the initial residual; p = z follows as a copy
__global__ void k_scalar_rho0(const double* __restrict__ part, DevState* st) {
  if (st->stopped) return;
  const double zr = wave_total(part, S_ZR);
  if (threadIdx.x == 0) {
    st->rho = zr;
    st->rho0 = zr;
    st->best_rho = zr;
    st->stag_count = 0;
    st->bestnr = 1.0;
    st->nr = 1.0;
    st->pending_copy = 0;
    st->itcnt = 0;
    st->t1 = wall_clock64();
    if (st->maxits == 0) {
      st->stopped = 1;
      st->stop_reason = 1;
    }
  }
}

__global__ void k_copy(int n, const double* __restrict__ src,
                       double* __restrict__ dst,
                       const DevState* __restrict__ st) {
  if (st->stopped) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// PCG iteration

// q = la * p, partial p.q (G lanes per row, grid-stride over rows)
template <int G>
__global__ void __launch_bounds__(kBlock)
k_spmv_dot(int n, const int* __restrict__ Lptr, const int* __restrict__ Lidx,
           const double* __restrict__ Lval, const double* __restrict__ p,
           double* __restrict__ q, double* __restrict__ part,
           const DevState* __restrict__ st) {
  if (st->stopped) return;
  const int lane = threadIdx.x % G;
  const int rows_per_grid = gridDim.x * (kBlock / G);
  double v[1] = {0.0};
  for (int row = blockIdx.x * (kBlock / G) + threadIdx.x / G; row < n;
       row += rows_per_grid) {
    const int k0 = Lptr[row], k1 = Lptr[row + 1];
    double acc = 0.0;
    for (int k = k0 + lane; k < k1; k += G) acc += Lval[k] * p[Lidx[k]];
    acc = group_sum<G>(acc);
    if (lane == 0) {
      q[row] = acc;
      v[0] += p[row] * acc;
    }
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) part[S_PQ * kRedBlocks + blockIdx.x] = v[0];
}

// itcnt += 1; al = rho / (p.q)   (src/pcg.jl:289-304)
__global__ void k_scalar_al(const double* __restrict__ part, DevState* st) {
  if (st->stopped) return;
  const double pq = wave_total(part, S_PQ);
  if (threadIdx.x == 0) {
    st->itcnt += 1;
    st->check10 = (st->itcnt % 10 == 0) ? 1 : 0;
    st->al = st->rho / pq;
  }
}

// x += al p; r -= al q; bestx <- x (previous iterate) if it improved;
// partials r.r and, every 10th iteration, sum|p|, sum|x|, sum|q|, sum|r|
// of the values before the update (src/pcg.jl:308-354, axpy2! :448-453).
__global__ void k_update(int n, double* __restrict__ x,
                         double* __restrict__ bestx,
                         const double* __restrict__ p,
                         const double* __restrict__ q, double* __restrict__ r,
                         double* __restrict__ part,
                         const DevState* __restrict__ st) {
  if (st->stopped) return;
  const double al = st->al;
  const bool chk = st->check10 != 0;
  const bool cp = st->pending_copy != 0;
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double xo = x[i], pi = p[i], qi = q[i], ro = r[i];
    if (cp) bestx[i] = xo;
    x[i] = xo + al * pi;
    const double rn = ro + (-al) * qi;
    r[i] = rn;
    v[0] += rn * rn;
    if (chk) {
      v[1] += fabs(pi);
      v[2] += fabs(xo);
      v[3] += fabs(qi);
      v[4] += fabs(ro);
    }
  }
  block_sum<5>(v);
  if (threadIdx.x == 0) {
    part[S_RR * kRedBlocks + blockIdx.x] = v[0];
    part[S_AP * kRedBlocks + blockIdx.x] = v[1];
    part[S_AX * kRedBlocks + blockIdx.x] = v[2];
    part[S_AQ * kRedBlocks + blockIdx.x] = v[3];
    part[S_AR * kRedBlocks + blockIdx.x] = v[4];
  }
}

// the two al stagnation checks, nr, bestnr and convergence
// (src/pcg.jl:308-361)
__global__ void k_scalar_nr(const double* __restrict__ part, DevState* st) {
  if (st->stopped) return;
  const double rr = wave_total(part, S_RR);
  const double sp = wave_total(part, S_AP);
  const double sx = wave_total(part, S_AX);
  const double sq = wave_total(part, S_AQ);
  const double sr = wave_total(part, S_AR);
  if (threadIdx.x == 0) {
    st->pending_copy = 0;  // k_update has consumed it
    if (st->check10) {
      if (st->al * sp < kEps * sx) {
        st->stopped = 1;
        st->stop_reason = 4;
        return;
      }
      if (st->al * sq < kEps * sr) {
        st->stopped = 1;
        st->stop_reason = 5;
        return;
      }
    }
    const double nr = sqrt(rr) / st->nb;
    st->nr = nr;
    if (nr < st->bestnr) {
      st->bestnr = nr;
      st->pending_copy = 1;
    }
    if (nr < st->tol) {
      st->stopped = 1;
      st->stop_reason = 0;
    }
  }
}

// rho, the stag_test counter, the inf checks, beta and its stagnation check,
// maxtime and maxits (src/pcg.jl:367-430)
__global__ void k_scalar_rho(const double* __restrict__ part, DevState* st) {
  if (st->stopped) return;
  const double zr = wave_total(part, S_ZR);
  const double sz = wave_total(part, S_AZ);
  const double sp = wave_total(part, S_AP2);
  if (threadIdx.x == 0) {
    const double oldrho = st->rho;
    const double rho = zr;
    st->rho = rho;
    if (rho < st->best_rho * st->stag_fac) {
      st->best_rho = rho;
      st->stag_count = 0;
    } else if (st->stag_test > 0) {
      if (st->best_rho > st->stag_fac * rho) {
        st->stag_count += 1;
        if (st->stag_count > st->stag_test) {
          st->stopped = 1;
          st->stop_reason = 3;
          return;
        }
      }
    }
    if (isinf(rho)) {
      st->stopped = 1;
      st->stop_reason = 7;
      return;
    }
    const double beta = rho / oldrho;
    st->beta = beta;
    if (isinf(beta)) {
      st->stopped = 1;
      st->stop_reason = 8;
      return;
    }
    if (st->check10) {
      if (beta * sp < kEps * sz) {
        st->stopped = 1;
        st->stop_reason = 6;
        return;
      }
    }
    // the reference tests maxtime after the p update; the update cannot
    // change bestx, so testing it here returns the same result
    if (st->has_maxtime && (wall_clock64() - st->t1) > st->maxticks) {
      st->stopped = 1;
      st->stop_reason = 2;
      return;
    }
    if (st->maxits >= 0 && st->itcnt >= st->maxits) {
      st->stopped = 1;
      st->stop_reason = 1;
    }
  }
}

// p = z + beta p  (bzbeta!, src/pcg.jl:456-461)
__global__ void k_pupdate(int n, const double* __restrict__ z,
                          double* __restrict__ p,
                          const DevState* __restrict__ st) {
  if (st->stopped) return;
  const double beta = st->beta;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    p[i] = z[i] + beta * p[i];
}

// bestx <- x when the last evaluated iteration improved (runs after stop)
__global__ void k_final_copy(int n, const double* __restrict__ x,
                             double* __restrict__ bestx,
                             const DevState* __restrict__ st) {
  if (!st->pending_copy) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    bestx[i] = x[i];
}

}  // namespace lapgpu
