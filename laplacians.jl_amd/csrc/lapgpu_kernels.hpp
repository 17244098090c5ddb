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

#include "lapgpu_device.hpp"

namespace lapgpu {
// ---------------------------------------------------------------------------
// The scalar logic of pcg (src/pcg.jl:256-430), one solve state at a time;
// the single-RHS kernels run it in thread 0, the multi-RHS kernels in thread
// j for right-hand side j.

// rho of the initial residual (src/pcg.jl:277-283)
__device__ inline void step_rho0(DevState* st, double zr) {
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

// itcnt += 1; al = rho / (p.q)   (src/pcg.jl:289-304)
__device__ inline void step_al(DevState* st, double pq) {
  st->itcnt += 1;
  st->check10 = (st->itcnt % 10 == 0) ? 1 : 0;
  st->al = st->rho / pq;
}

// the two al stagnation checks, nr, bestnr and convergence
// (src/pcg.jl:308-361)
__device__ inline void step_nr(DevState* st, double rr, double sp, double sx, double sq,
                               double sr) {
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
  st->bestnr_prev = st->bestnr;
  if (nr < st->bestnr) {
    st->bestnr = nr;
    st->pending_copy = 1;
  }
  if (st->fp32 && !st->spmv64 && nr < st->switch_tol) {
    // below switch_tol an FP32 SpMV cannot resolve the residual on badly
    // scaled graphs: continue with the FP64 SpMV, starting from an FP64
    // true residual
    st->spmv64 = 1;
    st->verify = 1;
  }
  if (nr < st->tol) {
    if (st->fp32) {
      st->verify = 1;  // confirmed by an FP64 true residual
    } else {
      st->stopped = 1;
      st->stop_reason = 0;
    }
  }
}

// rho, the stag_test counter, the inf checks, beta and its stagnation check,
// maxtime and maxits (src/pcg.jl:367-430)
__device__ inline void step_rho(DevState* st, double zr, double sz, double sp) {
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
        if (st->fp32 && !st->spmv64) {
          // FP32-inner mode: the FP32 SpMV's accuracy floor has been
          // reached above switch_tol; continue with the FP64 SpMV from an
          // FP64 true residual instead of stopping
          st->spmv64 = 1;
          st->verify = 1;
          st->stag_count = 0;
        } else {
          st->stopped = 1;
          st->stop_reason = 3;
          return;
        }
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
  if (threadIdx.x == 0) st->mu = st->bmu = st->center ? s / (double)n : 0.0;
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

// One group of consecutive levels (see LevelBlocks in lapgpu_plan.hpp):
//   out[row] = sum_k w[k] * (idx[k] < n ? A[idx[k]] : B[idx[k] - n])
// forward: A = y0, B = r, out = y0; backward: A = yF, B = y0, out = yF.
// Rows with more than kHeavy(G) entries are listed in `heavy` and handled
// by a whole wave each (threads [0, 64 nh)); the other rows get G lanes each.
// T is the sweep precision (double, or float in the FP32-inner mode).
// The sweeps (k_group, k_tail_ext) do not read the stop flag: they write
// only the sweep vectors, whose readers check it, and a flag read would
// put one more memory latency on each launch's critical path.
template <int G>
__host__ __device__ constexpr int heavy_limit() { return G >= 64 ? (1 << 30) : 8 * G; }

template <int G, typename T, typename TB>
__global__ void __launch_bounds__(kBlock)
k_group(int n, int lo, int hi, const int* __restrict__ heavy, int nh,
        const int* __restrict__ ptr, const int* __restrict__ idx,
        const T* __restrict__ w, const T* A, const TB* B, T* out) {
  const int gid = blockIdx.x * kBlock + threadIdx.x;
  if (gid < nh * kWave) {
    const int row = heavy[gid / kWave];
    const int lane = gid % kWave;
    const int k0 = ptr[row], k1 = ptr[row + 1];
    T acc = strided_dot<kWave, 8>(k0, k1, lane, idx, w,
                                  [&](int i) { return i < n ? A[i] : (T)B[i - n]; });
    acc = group_sum<kWave>(acc);
    if (lane == 0) out[row] = acc;
    return;
  }
  const int g2 = gid - nh * kWave;
  const int row = lo + g2 / G;
  const int lane = g2 % G;
  if (row >= hi) return;
  const int k0 = ptr[row], k1 = ptr[row + 1];
  if (k1 - k0 > heavy_limit<G>()) return;
  // at most 8 terms per lane: one round of loads
  T acc = strided_dot<G, 8>(k0, k1, lane, idx, w,
                            [&](int i) { return i < n ? A[i] : (T)B[i - n]; });
  acc = group_sum<G>(acc);
  if (lane == 0) out[row] = acc;
}

// ext[i] = r[T0+i] + sum of the forward in-edges of tail row i that come
// from outside the tail (G lanes per row)
template <int G, typename T>
__global__ void __launch_bounds__(kBlock)
k_tail_ext(int T0, int NT, const int* __restrict__ Eptr,
           const int* __restrict__ Eidx, const T* __restrict__ Ew,
           const double* __restrict__ r, const T* __restrict__ y,
           T* __restrict__ ext) {
  const int gid = blockIdx.x * kBlock + threadIdx.x;
  const int row = gid / G;
  const int lane = gid % G;
  if (row >= NT) return;
  const int k0 = Eptr[row], k1 = Eptr[row + 1];
  T acc = strided_dot<G, 8>(k0, k1, lane, Eidx, Ew, [&](int i) { return y[i]; });
  acc = group_sum<G>(acc);
  if (lane == 0) ext[row] = (T)r[T0 + row] + acc;
}

// forward + diagonal + backward on the tail; writes yF of the tail to y.
// E: entries per lane of the tail's slot programs.
template <typename T, typename S, int E>
__global__ void __launch_bounds__(kTailThreads)
k_tail(int T0, int NT, const T* __restrict__ ext, const T* __restrict__ s,
       T* __restrict__ yF, const TailStepD* __restrict__ fsteps, int nfsteps,
       const S* __restrict__ fslots, const TailStepD* __restrict__ bsteps,
       int nbsteps, const S* __restrict__ bslots,
       const DevState* __restrict__ st) {
  __shared__ T ly[kTailCapLds + kPartSlots];
  if (st->stopped) return;
  for (int i = threadIdx.x; i < NT; i += kTailThreads) ly[i] = ext[i];
  __syncthreads();
  run_program<tail_depth<E>(), true, E>(fsteps, nfsteps, fslots, ly);
  __syncthreads();
  for (int i = threadIdx.x; i < NT; i += kTailThreads) ly[i] *= s[T0 + i];
  __syncthreads();
  run_program<tail_depth<E>(), true, E>(bsteps, nbsteps, bslots, ly);
  __syncthreads();
  for (int i = threadIdx.x; i < NT; i += kTailThreads) yF[T0 + i] = ly[i];
}

// ---------------------------------------------------------------------------
// The tail's dense block [T1, n) (see DenseBlock in lapgpu_plan.hpp): its
// map from the forward inputs u to yF is the symmetric matrix D, stored
// row-major with leading dimension ldd (a multiple of 256, columns K..ldd
// zero).  A sweep with a dense block runs k_tail_fwd (the levels before T1,
// then the block's inputs), the product yF[T1..n) = D u, and k_tail_bwd (the
// levels before T1, backward).

// Column j = j0 + blockIdx.x of D: the unit vector e_j through the block's
// forward program, scaling and backward program, in LDS (the sweep's own
// arithmetic); stored as row j (D is symmetric).  Every vertex's mass
// reaches the root (the block's last vertex), so D holds s_root 1 1^T; that
// term only adds the constant s_root sum(u) to every yF (the backward rows
// are convex combinations, so it reaches every vertex unchanged), which the
// mean removal of LDLsolver cancels.  It is left out: it is D's largest part
// and cancels in D u, so D u keeps only the rounding of the rest.
template <int E>
__global__ void __launch_bounds__(kTailThreads)
k_dense_basis(int K, int T1, int j0, const double* __restrict__ s,
              const TailStepD* __restrict__ fsteps, int nfsteps,
              const TailSlotD* __restrict__ fslots, const TailStepD* __restrict__ bsteps,
              int nbsteps, const TailSlotD* __restrict__ bslots, double* __restrict__ D,
              long ldd) {
  __shared__ double ly[kTailCapLds + kPartSlots];
  const int j = j0 + blockIdx.x;
  if (j >= K) return;  // the whole workgroup
  for (int i = threadIdx.x; i < K; i += kTailThreads) ly[i] = i == j ? 1.0 : 0.0;
  __syncthreads();
  run_program<tail_depth<E>(), true, E>(fsteps, nfsteps, fslots, ly);
  __syncthreads();
  for (int i = threadIdx.x; i < K; i += kTailThreads) ly[i] *= s[T1 + i];
  __syncthreads();
  run_program<tail_depth<E>(), true, E>(bsteps, nbsteps, bslots, ly);
  __syncthreads();
  const double c = s[T1 + K - 1];  // the root's scale
  double* __restrict__ row = D + (size_t)j * ldd;
  for (int i = threadIdx.x; i < ldd; i += kTailThreads) row[i] = i < K ? ly[i] - c : 0.0;
}

// FP32 copy of D (FP32-inner mode)
__global__ void k_narrow(size_t count, const double* __restrict__ src, float* __restrict__ dst) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
    dst[i] = (float)src[i];
}

// Tail forward with a dense block: the levels of [T0, T1) and the block's
// inputs; writes s * y0 of [T0, T1) to yF and the block's inputs to u[0..K)
template <typename T, typename S, int E>
__global__ void __launch_bounds__(kTailThreads)
k_tail_fwd(int T0, int NT, int NT1, const T* __restrict__ ext, const T* __restrict__ s,
           T* __restrict__ yF, T* __restrict__ u, const TailStepD* __restrict__ fsteps,
           int nfsteps, const S* __restrict__ fslots, const DevState* __restrict__ st) {
  __shared__ T ly[kTailCapLds + kPartSlots];
  if (st->stopped) return;
  for (int i = threadIdx.x; i < NT; i += kTailThreads) ly[i] = ext[i];
  __syncthreads();
  run_program<tail_depth<E>(), true, E>(fsteps, nfsteps, fslots, ly);
  __syncthreads();
  for (int i = threadIdx.x; i < NT1; i += kTailThreads) yF[T0 + i] = ly[i] * s[T0 + i];
  for (int i = NT1 + threadIdx.x; i < NT; i += kTailThreads) u[i - NT1] = ly[i];
}

// Tail backward with a dense block: yF of [T0, T1) from its scaled y0 and
// the block's results yF[T1..n)
template <typename T, typename S, int E>
__global__ void __launch_bounds__(kTailThreads)
k_tail_bwd(int T0, int NT, int NT1, T* __restrict__ yF, const TailStepD* __restrict__ bsteps,
           int nbsteps, const S* __restrict__ bslots, const DevState* __restrict__ st) {
  __shared__ T ly[kTailCapLds + kPartSlots];
  if (st->stopped) return;
  for (int i = threadIdx.x; i < NT; i += kTailThreads) ly[i] = yF[T0 + i];
  __syncthreads();
  run_program<tail_depth<E>(), true, E>(bsteps, nbsteps, bslots, ly);
  __syncthreads();
  for (int i = threadIdx.x; i < NT1; i += kTailThreads) yF[T0 + i] = ly[i];
}

template <typename T>
struct Vec16;  // 16 bytes of T
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int n = 2;
  __device__ static double dot(const double2& a, const double2& b) { return a.x * b.x + a.y * b.y; }
};
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int n = 4;
  __device__ static double dot(const float4& a, const float4& b) {
    return (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
  }
};

// y[0..K) = D u (u zero beyond K up to ldd): kDenseRows rows per wave, each
// lane a strided share of 16-byte column chunks, summed in double in a fixed
// order (lane chunks in column order, then the wave's butterfly).
constexpr int kDenseRows = 8;
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_dense(int K, long ldd, const T* __restrict__ D, const T* __restrict__ u, T* __restrict__ y,
        const DevState* __restrict__ st) {
  if (st->stopped) return;
  using V = Vec16<T>;
  using VT = typename V::type;
  const int lane = threadIdx.x % kWave;
  const int r0 = ((blockIdx.x * kBlock + threadIdx.x) / kWave) * kDenseRows;
  if (r0 >= K) return;
  const int nr = min(kDenseRows, K - r0);
  double acc[kDenseRows];
#pragma unroll
  for (int r = 0; r < kDenseRows; ++r) acc[r] = 0.0;
  const VT* __restrict__ uv = reinterpret_cast<const VT*>(u);
  const long nv = ldd / V::n;
  for (long c = lane; c < nv; c += kWave) {
    const VT x = uv[c];
    VT d[kDenseRows];
#pragma unroll
    for (int r = 0; r < kDenseRows; ++r)
      if (r < nr) d[r] = reinterpret_cast<const VT*>(D + (size_t)(r0 + r) * ldd)[c];
#pragma unroll
    for (int r = 0; r < kDenseRows; ++r)
      if (r < nr) acc[r] += V::dot(d[r], x);
  }
#pragma unroll
  for (int r = 0; r < kDenseRows; ++r) {
    const double t = group_sum<kWave>(acc[r]);
    if (lane == 0 && r < nr) y[r0 + r] = (T)t;
  }
}

// partial sum of y (for the mean removal of LDLsolver)
template <typename T>
__global__ void k_vsum(int n, const T* __restrict__ y,
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
template <typename T>
__global__ void k_zdot(int n, const T* __restrict__ y,
                       const double* __restrict__ r,
                       const double* __restrict__ p, double* __restrict__ z,
                       double* __restrict__ part,
                       const DevState* __restrict__ st) {
  if (st->stopped) return;
  const double mu = st->mu;
  const bool chk = st->check10 != 0;
  double v[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double t = (double)y[i] - mu;
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

// FP32-inner mode SpMV: FP32 la and p until st->spmv64, FP64 afterwards
template <int G>
__global__ void __launch_bounds__(kBlock)
k_spmv_dot_mixed(int n, const int* __restrict__ Lptr, const int* __restrict__ Lidx,
                 const float* __restrict__ Lval32, const float* __restrict__ p32,
                 const double* __restrict__ Lval, const double* __restrict__ p,
                 double* __restrict__ q, double* __restrict__ part,
                 const DevState* __restrict__ st) {
  const int lb = xcd_block();
  if (st->stopped) return;
  const bool f64 = st->spmv64 != 0;
  const int lane = threadIdx.x % G;
  const int rows_per_grid = gridDim.x * (kBlock / G);
  const int first = lb * (kBlock / G) + threadIdx.x / G;
  double v[1] = {0.0};
  auto emit = [&](int row, double acc) {
    if (lane == 0) {
      q[row] = acc;
      v[0] += p[row] * acc;
    }
  };
  if (f64)
    rows_in_flight<G, 4>(first, rows_per_grid, n, Lptr, Lidx, Lval,
                         [&](int i) { return p[i]; }, emit);
  else
    rows_in_flight<G, 4>(first, rows_per_grid, n, Lptr, Lidx, Lval32,
                         [&](int i) { return p32[i]; },
                         [&](int row, float acc) { emit(row, (double)acc); });
  block_sum<1>(v);
  if (threadIdx.x == 0) part[S_PQ * kRedBlocks + lb] = v[0];
}

// rho of the initial residual; p = z follows as a copy
__global__ void k_scalar_rho0(const double* __restrict__ part, DevState* st) {
  if (st->stopped) return;
  const double zr = wave_total(part, S_ZR);
  if (threadIdx.x == 0) step_rho0(st, zr);
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
// T: matrix values and the p copy read (float in the FP32-inner mode);
// products are accumulated in T, q and p.q in double.
template <int G, typename T>
__global__ void __launch_bounds__(kBlock)
k_spmv_dot(int n, const int* __restrict__ Lptr, const int* __restrict__ Lidx,
           const T* __restrict__ Lval, const T* __restrict__ p,
           const double* __restrict__ pd, double* __restrict__ q,
           double* __restrict__ part, const DevState* __restrict__ st) {
  const int lb = xcd_block();
  if (st->stopped) return;
  const int lane = threadIdx.x % G;
  const int rows_per_grid = gridDim.x * (kBlock / G);
  double v[1] = {0.0};
  rows_in_flight<G, 4>(lb * (kBlock / G) + threadIdx.x / G, rows_per_grid, n, Lptr, Lidx,
                       Lval, [&](int i) { return p[i]; }, [&](int row, T acc) {
                         if (lane == 0) {
                           q[row] = (double)acc;
                           v[0] += pd[row] * (double)acc;
                         }
                       });
  block_sum<1>(v);
  if (threadIdx.x == 0) part[S_PQ * kRedBlocks + lb] = v[0];
}

// itcnt += 1; al = rho / (p.q)   (src/pcg.jl:289-304)
__global__ void k_scalar_al(const double* __restrict__ part, DevState* st) {
  if (st->stopped) return;
  const double pq = wave_total(part, S_PQ);
  if (threadIdx.x == 0) step_al(st, pq);
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
  if (threadIdx.x == 0) step_nr(st, rr, sp, sx, sq, sr);
}

// FP32-inner mode: r = r0 - L x in FP64 (r0 = b .- mean(b)), partial r.r.
// Runs when `always` (the periodic residual replacement) or when the
// recursive residual has just passed tol (st->verify).
template <int G>
__global__ void __launch_bounds__(kBlock)
k_resid(int n, int always, const int* __restrict__ Lptr,
        const int* __restrict__ Lidx, const double* __restrict__ Lval,
        const double* __restrict__ x, const double* __restrict__ r0,
        double* __restrict__ r, double* __restrict__ part,
        const DevState* __restrict__ st) {
  const int lb = xcd_block();
  if (st->stopped || !(always || st->verify)) return;
  const int lane = threadIdx.x % G;
  const int rows_per_grid = gridDim.x * (kBlock / G);
  double v[1] = {0.0};
  rows_in_flight<G, 4>(lb * (kBlock / G) + threadIdx.x / G, rows_per_grid, n, Lptr, Lidx,
                       Lval, [&](int i) { return x[i]; }, [&](int row, double acc) {
                         if (lane == 0) {
                           const double t = r0[row] - acc;
                           r[row] = t;
                           v[0] += t * t;
                         }
                       });
  block_sum<1>(v);
  if (threadIdx.x == 0) part[S_RR * kRedBlocks + lb] = v[0];
}

// After the solve: (b - bmu) - la x in FP64 for the returned x (level
// order), partial sums of squares
template <int G>
__global__ void __launch_bounds__(kBlock)
k_true_resid(int n, const int* __restrict__ Lptr, const int* __restrict__ Lidx,
             const double* __restrict__ Lval, const double* __restrict__ x,
             const double* __restrict__ b, double* __restrict__ part,
             const DevState* __restrict__ st) {
  const int lb = xcd_block();
  const int lane = threadIdx.x % G;
  const int rows_per_grid = gridDim.x * (kBlock / G);
  const double mu = st->bmu;
  double v[1] = {0.0};
  rows_in_flight<G, 4>(lb * (kBlock / G) + threadIdx.x / G, rows_per_grid, n, Lptr, Lidx,
                       Lval, [&](int i) { return x[i]; }, [&](int row, double acc) {
                         if (lane == 0) {
                           const double t = (b[row] - mu) - acc;
                           v[0] += t * t;
                         }
                       });
  block_sum<1>(v);
  if (threadIdx.x == 0) part[S_RR * kRedBlocks + lb] = v[0];
}

__global__ void k_scalar_true(const double* __restrict__ part, DevState* st) {
  const double rr = wave_total(part, S_RR);
  if (threadIdx.x == 0) st->true_nr = st->nb > 0.0 ? sqrt(rr) / st->nb : 0.0;
}

// After k_resid: the true relative residual replaces nr; convergence is
// accepted only on it (FP32-inner mode).
__global__ void k_scalar_resid(int always, const double* __restrict__ part, DevState* st) {
  if (st->stopped || !(always || st->verify)) return;
  const double rr = wave_total(part, S_RR);
  if (threadIdx.x == 0) {
    const double nr = sqrt(rr) / st->nb;
    st->nr = nr;
    // a true residual below switch_tol also ends the FP32 SpMV (the
    // recursive one may hover above it at the FP32 SpMV's accuracy floor)
    if (!st->spmv64 && nr < st->switch_tol) st->spmv64 = 1;
    // the true residual decides whether this iterate is the best so far
    st->bestnr = st->bestnr_prev;
    st->pending_copy = 0;
    if (nr < st->bestnr) {
      st->bestnr = nr;
      st->pending_copy = 1;
    }
    if (st->verify && nr < st->tol) {
      st->stopped = 1;
      st->stop_reason = 0;
    }
    st->verify = 0;
  }
}

// p = z + beta p with an FP32 copy of p for the FP32 SpMV
__global__ void k_pupdate_f(int n, const double* __restrict__ z,
                            double* __restrict__ p, float* __restrict__ pf,
                            const DevState* __restrict__ st) {
  if (st->stopped) return;
  const double beta = st->beta;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double t = z[i] + beta * p[i];
    p[i] = t;
    pf[i] = (float)t;
  }
}

// p = z (and its FP32 copy) after the setup's preconditioner
__global__ void k_copy_f(int n, const double* __restrict__ src,
                         double* __restrict__ dst, float* __restrict__ dstf,
                         const DevState* __restrict__ st) {
  if (st->stopped) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    dst[i] = src[i];
    dstf[i] = (float)src[i];
  }
}

// rho, the stag_test counter, the inf checks, beta and its stagnation check,
// maxtime and maxits (src/pcg.jl:367-430)
__global__ void k_scalar_rho(const double* __restrict__ part, DevState* st) {
  if (st->stopped) return;
  const double zr = wave_total(part, S_ZR);
  const double sz = wave_total(part, S_AZ);
  const double sp = wave_total(part, S_AP2);
  if (threadIdx.x == 0) step_rho(st, zr, sz, sp);
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
// Stop (reason 10, cancelled) the states j < k whose bit is set in mask
// and that have not stopped: every later kernel of the solve skips them,
// and the final copy returns their best iterate.
__global__ void k_cancel(DevState* st, int k, unsigned long long mask) {
  const int j = threadIdx.x;
  if (j < k && ((mask >> j) & 1ull) && !st[j].stopped) {
    st[j].stopped = 1;
    st[j].stop_reason = 10;
  }
}

__global__ void k_final_copy(int n, const double* __restrict__ x,
                             double* __restrict__ bestx,
                             const DevState* __restrict__ st) {
  if (!st->pending_copy) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    bestx[i] = x[i];
}

}  // namespace lapgpu
