// lapgpu_multi.hpp — k right-hand sides of one graph solved together.
//
// Each right-hand side runs its own pcg (src/pcg.jl:256-445) with its own
// scalars and stopping rules (state array mst[K], the same step_* logic as
// the single-RHS path); they share every launch.  Vectors are interleaved,
// element (position v, rhs j) at v*K + j with K a power of two in [8, 64]:
// the sparse structure (indices, weights) is read once per row for all K
// right-hand sides and every vector access is K contiguous doubles.
//
// flags[0] = 1 once every right-hand side has stopped (all kernels then
// return at once); flags[1] = this iteration is a 10th one (the stagnation
// checks need the 1-norm partials; all running right-hand sides share the
// iteration count, so one flag serves them all); FP32-inner mode:
// flags[3] = some running column waits for an FP64 true-residual check.
//
// FP32-inner mode (as the single-RHS path): SpMV and the LDLinv apply in
// FP32 (weights, y, z), everything else in FP64 — r, x, p, the dots,
// alpha/beta, and the true residual r = r0 - la x recomputed every k
// iterations and before a column is accepted as converged.
// Included by lapgpu.hip only.
#pragma once

#include "lapgpu_kernels.hpp"

namespace lapgpu {

constexpr int kMRed = 1024;  // fixed grid of the multi-RHS reduction kernels

// Block partial of NV sums per rhs j = threadIdx.x % K, written to
// part[(slot * kMRed + block) * K + j].
template <int K, int NV>
__device__ __forceinline__ void mblock_sum(double (&a)[NV], double* __restrict__ part,
                                           const int (&slots)[NV], int blk = -1) {
  if (blk < 0) blk = blockIdx.x;
  __shared__ double red[NV][kBlock / kWave][K];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
#pragma unroll
    for (int o = K; o < kWave; o <<= 1) a[i] += __shfl_xor(a[i], o, kWave);
  }
  if (lane < K) {
#pragma unroll
    for (int i = 0; i < NV; ++i) red[i][wid][lane] = a[i];
  }
  __syncthreads();
  if (threadIdx.x < K) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kBlock / kWave; ++w) s += red[i][w][threadIdx.x];
      part[((int64_t)slots[i] * kMRed + blk) * K + threadIdx.x] = s;
    }
  }
}

// Total of one slot for rhs j = threadIdx.x, valid in threads < K of a
// kScalarThreads block; fixed summation order.
template <int K>
__device__ __forceinline__ double mtotal(const double* __restrict__ part, int slot) {
  constexpr int SUB = kScalarThreads / K;
  __shared__ double red[SUB][K];
  const int j = threadIdx.x % K, sub = threadIdx.x / K;
  double s = 0.0;
#pragma unroll 4
  for (int b = sub; b < kMRed; b += SUB) s += part[((int64_t)slot * kMRed + b) * K + j];
  __syncthreads();
  red[sub][j] = s;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x < K) {
#pragma unroll 4
    for (int u = 0; u < SUB; ++u) t += red[u][threadIdx.x];
  }
  return t;
}

__device__ __forceinline__ double readlane_t(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float readlane_t(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <int K, typename T>
__device__ __forceinline__ void mbcast(int myi, T myw, int src, int& i, T& w) {
  if constexpr (K == kWave) {
    i = __builtin_amdgcn_readlane(myi, src);
    w = readlane_t(myw, src);
  } else {
    i = __shfl(myi, src, kWave);
    w = __shfl(myw, src, kWave);
  }
}

// acc + the sum over the entries [k0, k1) of a row of w[e] * src(idx[e]) for the K
// lanes (one per rhs) of row group `sub` of a wave, in entry order, in T.
// Indices and weights are loaded K at a time, one per lane, and broadcast
// within the group, so the gathers of one row are independent loads rather
// than a chain of index -> value loads; the gathers go out U at a time,
// the last (partial) batch predicated, so a row of up to U entries costs
// one gather latency (a row of 13-17 entries, the level groups' mean on an
// expander, used to wait for a batch of 8 and then for each remaining
// entry in turn).  For K = 64 the row is wave-uniform (k0, k1 must be
// uniform) and the broadcast is a readlane into scalar registers.
template <int K, typename T, typename F>
__device__ __forceinline__ T mrow_dot(int k0, int k1, int j, int sub, const int* __restrict__ idx,
                                      const T* __restrict__ w, F&& src, T acc = 0) {
  constexpr int U = K == kWave ? 16 : 8;  // gathers in flight per lane (U divides K)
  for (int base = k0; base < k1; base += K) {
    const int kk = base + j;
    int myi = 0;
    T myw = 0;
    if (kk < k1) {
      myi = idx[kk];
      myw = w[kk];
    }
    const int cnt = min(K, k1 - base);
    const int g0 = sub * K;
    for (int e = 0; e < cnt; e += U) {
      int i[U];
      T we[U], v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) mbcast<K>(myi, myw, g0 + e + u, i[u], we[u]);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (e + u < cnt) v[u] = src(i[u]);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (e + u < cnt) acc += we[u] * v[u];
    }
  }
  return acc;
}

template <int K>
__device__ __forceinline__ int wave_uniform(int v) {
  if constexpr (K == kWave) return __builtin_amdgcn_readfirstlane(v);
  return v;
}

// A grid-stride loop over rows for the K lanes of row group `sub` (rows
// first_row, first_row + stride, ...) with RU rows in flight: the row
// pointers of RU rows are loaded together, then each row's first K
// indices/weights, then the gathers of up to 8 entries of every row, before
// any of them is used.  Each row's sum is mrow_dot's (entries in order);
// emit(row, sum) is called for the rows in order.
template <int K, int RU, typename T, typename F, typename Emit>
__device__ __forceinline__ void mrows_in_flight(int first_row, int stride, int n, int j, int sub,
                                                const int* __restrict__ ptr,
                                                const int* __restrict__ idx,
                                                const T* __restrict__ w, F src, Emit emit) {
  constexpr int U = 8;
  const int g0 = sub * K;
  for (int r0 = first_row; r0 < n; r0 += RU * stride) {
    int k0[RU], k1[RU], myi[RU];
    T myw[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int row = r0 + u * stride;
      k0[u] = row < n ? wave_uniform<K>(ptr[row]) : 0;
      k1[u] = row < n ? wave_uniform<K>(ptr[row + 1]) : 0;
    }
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int kk = k0[u] + j;
      myi[u] = 0;
      myw[u] = 0;
      if (kk < k1[u]) {
        myi[u] = idx[kk];
        myw[u] = w[kk];
      }
    }
    int ii[RU][U];
    T we[RU][U], v[RU][U];
#pragma unroll
    for (int u = 0; u < RU; ++u)
#pragma unroll
      for (int e = 0; e < U; ++e) mbcast<K>(myi[u], myw[u], g0 + e, ii[u][e], we[u][e]);
#pragma unroll
    for (int u = 0; u < RU; ++u)
#pragma unroll
      for (int e = 0; e < U; ++e)
        if (e < k1[u] - k0[u]) v[u][e] = src(ii[u][e]);
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int row = r0 + u * stride;
      const int cnt = k1[u] - k0[u];
      T acc = 0;
#pragma unroll
      for (int e = 0; e < U; ++e)
        if (e < cnt) acc += we[u][e] * v[u][e];
      // entries U .. K-1 of the first chunk, then whole chunks (rare)
      for (int e = U; e < min(K, cnt); ++e) {
        int i;
        T wv;
        mbcast<K>(myi[u], myw[u], g0 + e, i, wv);
        acc += wv * src(i);
      }
      if (cnt > K) acc = mrow_dot<K, T>(k0[u] + K, k1[u], j, sub, idx, w, src, acc);
      if (row < n) emit(row, acc);
    }
  }
}

// after the per-rhs steps: flags[0] = every rhs stopped
template <int K>
__device__ __forceinline__ void mark_all(const DevState* mst, int* flags) {
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 1;
    for (int j = 0; j < K; ++j) s &= (mst[j].stopped != 0) ? 1 : 0;
    flags[0] = s;
  }
}

// ---- setup -----------------------------------------------------------------

// mb[v*K+j] = B[j*n + perm[v]] (zero for padding columns j >= k); sums
template <int K>
__global__ void __launch_bounds__(kBlock)
k_mgather(int n, int k, const int* __restrict__ perm, const double* __restrict__ B,
          double* __restrict__ mb, double* __restrict__ part) {
  const int j = threadIdx.x % K;
  const int64_t nK = (int64_t)n * K;
  double a[1] = {0.0};
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nK;
       e += (int64_t)gridDim.x * kBlock) {
    const int v = (int)(e / K);
    const double t = j < k ? B[(int64_t)j * n + perm[v]] : 0.0;
    mb[e] = t;
    a[0] += t;
  }
  const int slots[1] = {S_SUM};
  mblock_sum<K, 1>(a, part, slots);
}

template <int K>
__global__ void k_mscalar_mean(int n, const double* __restrict__ part, DevState* mst) {
  const double s = mtotal<K>(part, S_SUM);
  if (threadIdx.x < K) {
    DevState* st = mst + threadIdx.x;
    st->mu = st->bmu = st->center ? s / (double)n : 0.0;
  }
}

// r = b - mean(b); x = bestx = 0; partial r.r
template <int K>
__global__ void __launch_bounds__(kBlock)
k_mcenter(int n, const double* __restrict__ mb, double* __restrict__ r, double* __restrict__ x,
          double* __restrict__ bestx, double* __restrict__ part,
          const DevState* __restrict__ mst) {
  const int j = threadIdx.x % K;
  const double mu = mst[j].mu;
  const int64_t nK = (int64_t)n * K;
  double a[1] = {0.0};
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nK;
       e += (int64_t)gridDim.x * kBlock) {
    const double t = mb[e] - mu;
    r[e] = t;
    x[e] = 0.0;
    bestx[e] = 0.0;
    a[0] += t * t;
  }
  const int slots[1] = {S_RR};
  mblock_sum<K, 1>(a, part, slots);
}

template <int K>
__global__ void k_mscalar_nb(const double* __restrict__ part, DevState* mst, int* flags) {
  const double rr = mtotal<K>(part, S_RR);
  if (threadIdx.x < K) {
    DevState* st = mst + threadIdx.x;
    st->nb = sqrt(rr);
    if (st->nb == 0.0) {
      st->stopped = 1;
      st->stop_reason = 9;  // zero rhs (and every padding column)
    }
  }
  mark_all<K>(mst, flags);
}

// ---- the LDLinv apply --------------------------------------------------------

// One level group (composite rows, as k_group) for K right-hand sides:
//   out[v*K+j] = sum_e w[e] * (idx[e] < n ? A[idx[e]*K+j] : B[(idx[e]-n)*K+j])
// forward: A = y0, B = r, out = y0; backward: A = yF, B = y0, out = yF.
//  * rows longer than kMHeavyLen (heavy[0..nh)): one kMGroupThreads-thread
//    workgroup each; thread (slice s, rhs j) sums the entries s, s + SL, ...
//    (SL = kMGroupThreads / K) and the slices are added in a fixed tree
//    order in LDS (mgroup_heavy_row);
//  * the light rows lo..hi: a wave holds 64 / (K S) rows; lane (row r,
//    slice s, rhs j) sums the entries s, s + S, ... of row r, and the S
//    slices are added with shuffles (S = 1: one lane per rhs walks the
//    whole row, indices broadcast, see mrow_dot).
// For K <= 32 both kinds share one launch of kMGroupThreads-thread
// workgroups, the heavy rows first (a launch costs a few microseconds, and
// on the expanders most groups have rows of both kinds; 64-column rows ran
// slower in workgroups of that size, so there the heavy rows keep a launch
// of their own, k_mgroup_heavy, after one of kBlock-thread workgroups).
// The sums do not depend on the split: the same bits either way.
// The stop flag is loaded together with the first index loads.
constexpr int kMHeavyLen = 64;
constexpr int kMGroupThreads = 1024;
// entries whose loads a lane of a sliced row issues together (strided_dot):
// a slice of up to kMSliceU entries waits for one index and one gather
// latency (the sums are added in entry order whatever the batch).  Measured
// against batches of 4 (FP32-inner LDLinv apply, 3.4M-vertex expander):
// K = 8 / 16 / 32 / 64 -9 / -12 / -15 / -16 %; batches of 16 were slower
// for K <= 32 (registers), faster for the FP32 64-column long rows (-4 %)
constexpr int kMSliceU = 8;
template <int K, typename T>
constexpr int heavy_slice_u() { return K == kWave && sizeof(T) == 4 ? 16 : kMSliceU; }

template <int K, typename T, typename F>
__device__ __forceinline__ void mgroup_heavy_row(int row, const int* __restrict__ ptr,
                                                 const int* __restrict__ idx,
                                                 const T* __restrict__ w, F&& src, T* out) {
  constexpr int SL = kMGroupThreads / K;
  __shared__ T red[SL][K];
  const int j = threadIdx.x % K, s = threadIdx.x / K;
  const int k0 = ptr[row], k1 = ptr[row + 1];
  red[s][j] = strided_dot<SL, heavy_slice_u<K, T>()>(k0, k1, s, idx, w, src);
  __syncthreads();
#pragma unroll
  for (int h = SL / 2; h > 0; h >>= 1) {
    if (s < h) red[s][j] += red[s + h][j];
    __syncthreads();
  }
  if (s == 0) out[(int64_t)row * K + j] = red[0][j];
}

// NT-thread workgroups; the first nh take the heavy rows (NT must then be
// kMGroupThreads), the rest the light rows.
template <int K, int S, int NT, typename T, typename TB>
__global__ void __launch_bounds__(NT)
k_mgroup(int n, int lo, int hi, const int* __restrict__ heavy, int nh,
         const int* __restrict__ ptr, const int* __restrict__ idx, const T* __restrict__ w,
         const T* A, const TB* B, T* out, const int* __restrict__ flags) {
  const int j = threadIdx.x % K;
  auto src = [&](int i) -> T {
    return i < n ? A[(int64_t)i * K + j] : (T)B[(int64_t)(i - n) * K + j];
  };
  if constexpr (NT == kMGroupThreads) {
    if ((int)blockIdx.x < nh) {
      const int stop = flags[0];
      const int row = heavy[blockIdx.x];
      if (stop) return;
      mgroup_heavy_row<K, T>(row, ptr, idx, w, src, out);
      return;
    }
  }
  constexpr int R = kWave / (K * S);  // rows per wave
  const int lane = threadIdx.x % kWave;
  const int s = (lane / K) % S, r = lane / (K * S);
  const int row = lo + (((int)blockIdx.x - nh) * NT + (int)threadIdx.x) / kWave * R + r;
  const int stop = flags[0];
  int k0 = 0, k1 = 0;
  if (row < hi) {
    k0 = ptr[row];
    k1 = ptr[row + 1];
  }
  if (stop) return;
  if constexpr (S == 1) {
    if (row >= hi) return;  // whole row groups leave together
    k0 = wave_uniform<K>(k0);
    k1 = wave_uniform<K>(k1);
    if (k1 - k0 > kMHeavyLen) return;
    out[(int64_t)row * K + j] = mrow_dot<K, T>(k0, k1, j, r, idx, w, src);
  } else {
    T acc = 0;
    bool mine = false;
    if (row < hi && k1 - k0 <= kMHeavyLen) {
      mine = true;
      acc = strided_dot<S, kMSliceU>(k0, k1, s, idx, w, src);
    }
#pragma unroll
    for (int o = K; o < K * S; o <<= 1) acc += __shfl_xor(acc, o, kWave);
    if (mine && s == 0) out[(int64_t)row * K + j] = acc;
  }
}

// The heavy rows of a group in a launch of their own (heavy[blockIdx.x])
template <int K, typename T, typename TB>
__global__ void __launch_bounds__(kMGroupThreads)
k_mgroup_heavy(int n, const int* __restrict__ heavy, const int* __restrict__ ptr,
               const int* __restrict__ idx, const T* __restrict__ w, const T* A, const TB* B,
               T* out, const int* __restrict__ flags) {
  const int j = threadIdx.x % K;
  auto src = [&](int i) -> T {
    return i < n ? A[(int64_t)i * K + j] : (T)B[(int64_t)(i - n) * K + j];
  };
  const int stop = flags[0];
  const int row = heavy[blockIdx.x];
  if (stop) return;
  mgroup_heavy_row<K, T>(row, ptr, idx, w, src, out);
}

// ext[i*K+j] = r[(T0+i)*K+j] + the forward in-edges of tail row i from
// outside the tail (as k_tail_ext), one lane per (row, rhs)
template <int K, typename T>
__global__ void __launch_bounds__(kBlock)
k_mtail_ext(int T0, int NT, const int* __restrict__ Eptr, const int* __restrict__ Eidx,
            const T* __restrict__ Ew, const double* __restrict__ r, const T* __restrict__ y,
            T* __restrict__ ext, const int* __restrict__ flags) {
  if (flags[0]) return;
  constexpr int R = kWave / K;
  const int lane = threadIdx.x % kWave, j = lane % K, sub = lane / K;
  const int row = ((blockIdx.x * kBlock + threadIdx.x) / kWave) * R + sub;
  if (row >= NT) return;
  const int k0 = wave_uniform<K>(Eptr[row]), k1 = wave_uniform<K>(Eptr[row + 1]);
  const T acc =
      mrow_dot<K, T>(k0, k1, j, sub, Eidx, Ew, [&](int i) -> T { return y[(int64_t)i * K + j]; });
  ext[(int64_t)row * K + j] = (T)r[(int64_t)(T0 + row) * K + j] + acc;
}

// The tail for rhs j = blockIdx.x: the single-RHS slot programs on an
// LDS-resident copy of column j (as k_tail); K workgroups run side by side
// on K CUs.  Columns that have stopped are skipped.
template <int K, typename T, typename S, int E>
__global__ void __launch_bounds__(kTailThreads)
k_mtail(int T0, int NT, const T* __restrict__ ext, const T* __restrict__ s, T* __restrict__ z,
        const TailStepD* __restrict__ fsteps, int nfsteps, const S* __restrict__ fslots,
        const TailStepD* __restrict__ bsteps, int nbsteps, const S* __restrict__ bslots,
        const DevState* __restrict__ mst, const int* __restrict__ flags) {
  __shared__ T ly[kTailCapLds + kPartSlots];
  const int j = blockIdx.x;
  if (flags[0] || mst[j].stopped) return;
  for (int i = threadIdx.x; i < NT; i += kTailThreads) ly[i] = ext[(int64_t)i * K + j];
  __syncthreads();
  run_program<tail_depth<E>(), true, E>(fsteps, nfsteps, fslots, ly);
  __syncthreads();
  for (int i = threadIdx.x; i < NT; i += kTailThreads) ly[i] *= s[T0 + i];
  __syncthreads();
  run_program<tail_depth<E>(), true, E>(bsteps, nbsteps, bslots, ly);
  __syncthreads();
  for (int i = threadIdx.x; i < NT; i += kTailThreads) z[(int64_t)(T0 + i) * K + j] = ly[i];
}

// With a dense block (k_tail_fwd / k_tail_bwd for K columns): the tail
// forward for column j = blockIdx.x writes s * y0 of [T0, T1) to z and the
// block's inputs to u (row i of the block at u[i*K + j]) ...
template <int K, typename T, typename S, int E>
__global__ void __launch_bounds__(kTailThreads)
k_mtail_fwd(int T0, int NT, int NT1, const T* __restrict__ ext, const T* __restrict__ s,
            T* __restrict__ z, T* __restrict__ u, const TailStepD* __restrict__ fsteps,
            int nfsteps, const S* __restrict__ fslots, const DevState* __restrict__ mst,
            const int* __restrict__ flags) {
  __shared__ T ly[kTailCapLds + kPartSlots];
  const int j = blockIdx.x;
  if (flags[0] || mst[j].stopped) return;
  for (int i = threadIdx.x; i < NT; i += kTailThreads) ly[i] = ext[(int64_t)i * K + j];
  __syncthreads();
  run_program<tail_depth<E>(), true, E>(fsteps, nfsteps, fslots, ly);
  __syncthreads();
  for (int i = threadIdx.x; i < NT1; i += kTailThreads)
    z[(int64_t)(T0 + i) * K + j] = ly[i] * s[T0 + i];
  for (int i = NT1 + threadIdx.x; i < NT; i += kTailThreads) u[(int64_t)(i - NT1) * K + j] = ly[i];
}

// ... and the backward reads them back with the block's results
template <int K, typename T, typename S, int E>
__global__ void __launch_bounds__(kTailThreads)
k_mtail_bwd(int T0, int NT, int NT1, T* __restrict__ z, const TailStepD* __restrict__ bsteps,
            int nbsteps, const S* __restrict__ bslots, const DevState* __restrict__ mst,
            const int* __restrict__ flags) {
  __shared__ T ly[kTailCapLds + kPartSlots];
  const int j = blockIdx.x;
  if (flags[0] || mst[j].stopped) return;
  for (int i = threadIdx.x; i < NT; i += kTailThreads) ly[i] = z[(int64_t)(T0 + i) * K + j];
  __syncthreads();
  run_program<tail_depth<E>(), true, E>(bsteps, nbsteps, bslots, ly);
  __syncthreads();
  for (int i = threadIdx.x; i < NT1; i += kTailThreads) z[(int64_t)(T0 + i) * K + j] = ly[i];
}

// The dense block for K columns on the matrix cores: Y = D U with D (Kd x
// Kd, leading dimension ldd) and U (rows of K, zero from row Kd to ldd).
// Workgroup (row tile rt, k split ks) computes rows [64 rt, 64 rt + 64) over
// columns [ks len, ks len + len) of D into partial ks: each of its four
// waves owns 16 rows and holds K/16 (at least one) 16x16 accumulator tiles;
// per 64-column slab the slab's U rows are staged in LDS, and each 4-wide
// k step is one 16x16x4 MFMA per tile (A: D[row l%16][k + l/16] from global,
// B: U[k + l/16][col l%16] from LDS).  The partials are added in split order
// by k_mdense_sum, so the result does not depend on scheduling.
typedef double dense_d4 __attribute__((ext_vector_type(4)));
typedef float dense_f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = dense_d4;
  // output item v of lane l: row l/16 + 4 v, column l%16
  __device__ static int row(int l, int v) { return l / 16 + 4 * v; }
  __device__ static acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  using acc_t = dense_f4;
  // output item v of lane l: row 4 (l/16) + v, column l%16
  __device__ static int row(int l, int v) { return 4 * (l / 16) + v; }
  __device__ static acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

constexpr int kMDenseRows = 64;   // rows per workgroup (16 per wave)
constexpr int kMDenseSlab = 64;   // columns of D per LDS stage

template <int K, typename T>
__global__ void __launch_bounds__(kBlock)
k_mdense(int Kd, long ldd, const T* __restrict__ D, const T* __restrict__ U,
         T* __restrict__ part, long len, const int* __restrict__ flags) {
  if (flags[0]) return;
  constexpr int NTILE = K >= 16 ? K / 16 : 1;
  constexpr int KC = NTILE * 16;  // columns covered by the tiles (K = 8: 16, the upper half zero)
  __shared__ T us[kMDenseSlab][KC];
  using M = Mfma<T>;
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const int r0 = blockIdx.x * kMDenseRows + w * 16;
  const long kb = (long)blockIdx.y * len, ke = min(kb + len, ldd);
  typename M::acc_t acc[NTILE];
#pragma unroll
  for (int t = 0; t < NTILE; ++t) acc[t] = typename M::acc_t{0, 0, 0, 0};
  const int arow = r0 + lane % 16;
  const bool arow_ok = arow < Kd;
  const T* __restrict__ Drow = D + (size_t)(arow_ok ? arow : 0) * ldd;
  for (long k0 = kb; k0 < ke; k0 += kMDenseSlab) {
    // stage U rows [k0, k0 + 64) (ldd is a multiple of 64: whole slabs)
    for (int e = threadIdx.x; e < kMDenseSlab * KC; e += kBlock) {
      const int kk = e / KC, col = e % KC;
      us[kk][col] = col < K ? U[(size_t)(k0 + kk) * K + col] : (T)0;
    }
    __syncthreads();
#pragma unroll 4
    for (int s = 0; s < kMDenseSlab / 4; ++s) {
      const int kk = 4 * s + lane / 16;
      const T a = arow_ok ? Drow[k0 + kk] : (T)0;
#pragma unroll
      for (int t = 0; t < NTILE; ++t) acc[t] = M::step(a, us[kk][16 * t + lane % 16], acc[t]);
    }
    __syncthreads();
  }
  T* __restrict__ P = part + (size_t)blockIdx.y * Kd * K;
#pragma unroll
  for (int t = 0; t < NTILE; ++t) {
    const int col = 16 * t + lane % 16;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = r0 + M::row(lane, v);
      if (row < Kd && col < K) P[(size_t)row * K + col] = acc[t][v];
    }
  }
}

// y[i*K + j] = sum of the nsplit partials, in split order
template <int K, typename T>
__global__ void __launch_bounds__(kBlock)
k_mdense_sum(int Kd, int nsplit, const T* __restrict__ part, T* __restrict__ y,
             const int* __restrict__ flags) {
  if (flags[0]) return;
  const size_t cnt = (size_t)Kd * K;
  for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < cnt; e += (size_t)gridDim.x * kBlock) {
    T acc = part[e];
    for (int s = 1; s < nsplit; ++s) acc += part[(size_t)s * cnt + e];
    y[e] = acc;
  }
}

// partial sums of yF (mean removal of LDLsolver)
template <int K, typename T>
__global__ void __launch_bounds__(kBlock)
k_mvsum(int n, const T* __restrict__ z, double* __restrict__ part,
        const int* __restrict__ flags) {
  if (flags[0]) return;
  const int64_t nK = (int64_t)n * K;
  const int64_t S = (int64_t)gridDim.x * kBlock;
  double a[1] = {0.0};
  int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  // four grid strides per pass with their loads issued together; the sums
  // run in the same order as one stride at a time
  for (; e + 3 * S < nK; e += 4 * S) {
    T v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = z[e + u * S];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[0] += (double)v[u];
  }
  for (; e < nK; e += S) a[0] += (double)z[e];
  const int slots[1] = {S_SUM};
  mblock_sum<K, 1>(a, part, slots);
}

template <int K>
__global__ void k_mscalar_mu(int n, const double* __restrict__ part, DevState* mst,
                             const int* __restrict__ flags) {
  if (flags[0]) return;
  const double s = mtotal<K>(part, S_SUM);
  if (threadIdx.x < K) mst[threadIdx.x].mu = s / (double)n;
}

// partials (yF - mu).r and, on 10th iterations, sum|yF - mu|, sum|p|
template <int K, typename T>
__global__ void __launch_bounds__(kBlock)
k_mzdot(int n, const T* __restrict__ z, const double* __restrict__ r,
        const double* __restrict__ p, double* __restrict__ part,
        const DevState* __restrict__ mst, const int* __restrict__ flags) {
  if (flags[0]) return;
  const bool chk = flags[1] != 0;
  const int j = threadIdx.x % K;
  const double mu = mst[j].mu;
  const int64_t nK = (int64_t)n * K;
  const int64_t S = (int64_t)gridDim.x * kBlock;
  double a[3] = {0.0, 0.0, 0.0};
  auto one = [&](double zv, double rv, int64_t e) {
    const double t = zv - mu;
    a[0] += t * rv;
    if (chk) {
      a[1] += fabs(t);
      a[2] += fabs(p[e]);
    }
  };
  int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  for (; e + 3 * S < nK; e += 4 * S) {
    double zv[4], rv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      zv[u] = (double)z[e + u * S];
      rv[u] = r[e + u * S];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) one(zv[u], rv[u], e + u * S);
  }
  for (; e < nK; e += S) one((double)z[e], r[e], e);
  const int slots[3] = {S_ZR, S_AZ, S_AP2};
  mblock_sum<K, 3>(a, part, slots);
}

template <int K>
__global__ void k_mscalar_rho0(const double* __restrict__ part, DevState* mst, int* flags) {
  if (flags[0]) return;
  const double zr = mtotal<K>(part, S_ZR);
  if (threadIdx.x < K && !mst[threadIdx.x].stopped) step_rho0(mst + threadIdx.x, zr);
  mark_all<K>(mst, flags);
}

// p = yF - mu after the setup's preconditioner (and its FP32 copy when pf)
template <int K, typename T>
__global__ void __launch_bounds__(kBlock)
k_mpinit(int n, const T* __restrict__ z, double* __restrict__ p, float* __restrict__ pf,
         const DevState* __restrict__ mst, const int* __restrict__ flags) {
  if (flags[0]) return;
  const int j = threadIdx.x % K;
  const double mu = mst[j].mu;
  const int64_t nK = (int64_t)n * K;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nK;
       e += (int64_t)gridDim.x * kBlock) {
    const double t = (double)z[e] - mu;
    p[e] = t;
    if (pf) pf[e] = (float)t;
  }
}

// ---- one PCG iteration ---------------------------------------------------------

// q = la * p, partial p.q; K lanes per row, one per rhs
template <int K>
__global__ void __launch_bounds__(kBlock)
k_mspmv(int n, const int* __restrict__ Lptr, const int* __restrict__ Lidx,
        const double* __restrict__ Lval, const double* __restrict__ p, double* __restrict__ q,
        double* __restrict__ part, const int* __restrict__ flags) {
  if (flags[0]) return;
  constexpr int R = kWave / K;  // rows per wave
  const int lane = threadIdx.x % kWave, j = lane % K, sub = lane / K;
  const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int nwaves = gridDim.x * (kBlock / kWave);
  double a[1] = {0.0};
  mrows_in_flight<K, 2>(wave * R + sub, nwaves * R, n, j, sub, Lptr, Lidx, Lval,
                        [&](int i) { return p[(int64_t)i * K + j]; }, [&](int row, double acc) {
                          q[(int64_t)row * K + j] = acc;
                          a[0] += p[(int64_t)row * K + j] * acc;
                        });
  const int slots[1] = {S_PQ};
  mblock_sum<K, 1>(a, part, slots);
}

// FP32-inner mode: q = la * p per column from the FP32 copies of la and p
// (sums in FP32) until that column leaves the FP32 SpMV (its spmv64, as the
// single-RHS k_spmv_dot_mixed), from the FP64 ones afterwards; p.q in FP64.
// Both weight arrays are broadcast; each lane gathers from its column's
// precision.
template <int K>
__global__ void __launch_bounds__(kBlock)
k_mspmv_mixed(int n, const int* __restrict__ Lptr, const int* __restrict__ Lidx,
              const float* __restrict__ Lval32, const float* __restrict__ pf,
              const double* __restrict__ Lval, const double* __restrict__ p,
              double* __restrict__ q, double* __restrict__ part,
              const DevState* __restrict__ mst, const int* __restrict__ flags) {
  if (flags[0]) return;
  constexpr int R = kWave / K;
  const int lane = threadIdx.x % kWave, j = lane % K, sub = lane / K;
  const bool f64 = mst[j].spmv64 != 0;
  const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int nwaves = gridDim.x * (kBlock / kWave);
  double a[1] = {0.0};
  // RU rows in flight (as mrows_in_flight): pointers, first K indices and
  // both weights, then up to 8 gathers per row, each lane from its
  // column's precision; sums in entry order, in that precision
  constexpr int RU = 2, U = 8;
  const int stride = nwaves * R, g0 = sub * K;
  for (int r0 = wave * R + sub; r0 < n; r0 += RU * stride) {
    int k0[RU], k1[RU], myi[RU];
    double myw[RU];
    float myw32[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int row = r0 + u * stride;
      k0[u] = row < n ? wave_uniform<K>(Lptr[row]) : 0;
      k1[u] = row < n ? wave_uniform<K>(Lptr[row + 1]) : 0;
    }
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int kk = k0[u] + j;
      myi[u] = 0;
      myw[u] = 0.0;
      myw32[u] = 0.0f;
      if (kk < k1[u]) {
        myi[u] = Lidx[kk];
        myw[u] = Lval[kk];
        myw32[u] = Lval32[kk];
      }
    }
    int ii[RU][U];
    double we[RU][U], v64[RU][U];
    float we32[RU][U], v32[RU][U];
#pragma unroll
    for (int u = 0; u < RU; ++u)
#pragma unroll
      for (int e = 0; e < U; ++e) {
        int i2;
        mbcast<K>(myi[u], myw[u], g0 + e, ii[u][e], we[u][e]);
        mbcast<K>(myi[u], myw32[u], g0 + e, i2, we32[u][e]);
      }
#pragma unroll
    for (int u = 0; u < RU; ++u)
#pragma unroll
      for (int e = 0; e < U; ++e)
        if (e < k1[u] - k0[u]) {
          if (f64)
            v64[u][e] = p[(int64_t)ii[u][e] * K + j];
          else
            v32[u][e] = pf[(int64_t)ii[u][e] * K + j];
        }
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int row = r0 + u * stride;
      const int cnt = k1[u] - k0[u];
      double a64 = 0.0;
      float a32 = 0.0f;
#pragma unroll
      for (int e = 0; e < U; ++e)
        if (e < cnt) {
          if (f64)
            a64 += we[u][e] * v64[u][e];
          else
            a32 += we32[u][e] * v32[u][e];
        }
      // the rest of the row (longer rows; rare)
      for (int base = k0[u]; base < k1[u]; base += K) {
        int mi = myi[u];
        double mw = myw[u];
        float mw32 = myw32[u];
        if (base != k0[u]) {
          const int kk = base + j;
          mi = 0;
          mw = 0.0;
          mw32 = 0.0f;
          if (kk < k1[u]) {
            mi = Lidx[kk];
            mw = Lval[kk];
            mw32 = Lval32[kk];
          }
        }
        const int c = min(K, k1[u] - base);
        for (int e = base == k0[u] ? U : 0; e < c; ++e) {
          int i, i2;
          double w;
          float w32;
          mbcast<K>(mi, mw, g0 + e, i, w);
          mbcast<K>(mi, mw32, g0 + e, i2, w32);
          if (f64)
            a64 += w * p[(int64_t)i * K + j];
          else
            a32 += w32 * pf[(int64_t)i * K + j];
        }
      }
      if (row < n) {
        const double acc = f64 ? a64 : (double)a32;
        q[(int64_t)row * K + j] = acc;
        a[0] += p[(int64_t)row * K + j] * acc;
      }
    }
  }
  const int slots[1] = {S_PQ};
  mblock_sum<K, 1>(a, part, slots);
}

// itcnt += 1; al = rho / (p.q); flags[1] = 10th iteration
template <int K>
__global__ void k_mscalar_al(const double* __restrict__ part, DevState* mst, int* flags) {
  if (flags[0]) return;
  const double pq = mtotal<K>(part, S_PQ);
  if (threadIdx.x < K && !mst[threadIdx.x].stopped) step_al(mst + threadIdx.x, pq);
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0;
    for (int j = 0; j < K; ++j)
      if (!mst[j].stopped) c |= mst[j].check10;
    flags[1] = c;
  }
}

// x += al p; r -= al q; bestx <- x where the previous iterate improved;
// partials r.r and (10th iterations) sum|p|, sum|x|, sum|q|, sum|r| of the
// values before the update (src/pcg.jl:308-354).  Stopped rhs are frozen.
template <int K>
__global__ void __launch_bounds__(kBlock)
k_mupdate(int n, double* __restrict__ x, double* __restrict__ bestx,
          const double* __restrict__ p, const double* __restrict__ q, double* __restrict__ r,
          double* __restrict__ part, const DevState* __restrict__ mst,
          const int* __restrict__ flags) {
  if (flags[0]) return;
  const bool chk = flags[1] != 0;
  const int j = threadIdx.x % K;
  const bool run = mst[j].stopped == 0;
  const bool cp = mst[j].pending_copy != 0;
  const double al = mst[j].al;
  const int64_t nK = (int64_t)n * K;
  double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (run) {
    auto one = [&](int64_t e, double xo, double pi, double qi, double ro) {
      if (cp) bestx[e] = xo;
      x[e] = xo + al * pi;
      const double rn = ro + (-al) * qi;
      r[e] = rn;
      a[0] += rn * rn;
      if (chk) {
        a[1] += fabs(pi);
        a[2] += fabs(xo);
        a[3] += fabs(qi);
        a[4] += fabs(ro);
      }
    };
    // four grid strides per pass with their loads issued together; the
    // sums run in the same order as one stride at a time
    const int64_t S = (int64_t)gridDim.x * kBlock;
    int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    for (; e + 3 * S < nK; e += 4 * S) {
      double xo[4], pi[4], qi[4], ro[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        xo[u] = x[e + u * S];
        pi[u] = p[e + u * S];
        qi[u] = q[e + u * S];
        ro[u] = r[e + u * S];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) one(e + u * S, xo[u], pi[u], qi[u], ro[u]);
    }
    for (; e < nK; e += S) one(e, x[e], p[e], q[e], r[e]);
  }
  const int slots[5] = {S_RR, S_AP, S_AX, S_AQ, S_AR};
  mblock_sum<K, 5>(a, part, slots);
}

template <int K>
__global__ void k_mscalar_nr(const double* __restrict__ part, DevState* mst, int* flags) {
  if (flags[0]) return;
  const double rr = mtotal<K>(part, S_RR);
  const double sp = mtotal<K>(part, S_AP);
  const double sx = mtotal<K>(part, S_AX);
  const double sq = mtotal<K>(part, S_AQ);
  const double sr = mtotal<K>(part, S_AR);
  if (threadIdx.x < K && !mst[threadIdx.x].stopped)
    step_nr(mst + threadIdx.x, rr, sp, sx, sq, sr);
  mark_all<K>(mst, flags);
  if (threadIdx.x == 0) {
    int anyv = 0;
    for (int j = 0; j < K; ++j)
      if (!mst[j].stopped) anyv |= mst[j].verify;
    flags[3] = anyv;
  }
}

// FP32-inner mode: r = r0 - la x in FP64 for the running columns, when
// `always` (the periodic residual replacement) or for the columns whose
// recursive residual has just passed tol (verify); partial r.r
template <int K>
__global__ void __launch_bounds__(kBlock)
k_mresid(int n, int always, const int* __restrict__ Lptr, const int* __restrict__ Lidx,
         const double* __restrict__ Lval, const double* __restrict__ x,
         const double* __restrict__ r0, double* __restrict__ r, double* __restrict__ part,
         const DevState* __restrict__ mst, const int* __restrict__ flags) {
  if (flags[0] || !(always || flags[3])) return;
  constexpr int R = kWave / K;
  const int lane = threadIdx.x % kWave, j = lane % K, sub = lane / K;
  const bool run = !mst[j].stopped && (always || mst[j].verify);
  const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int nwaves = gridDim.x * (kBlock / kWave);
  double a[1] = {0.0};
  mrows_in_flight<K, 2>(wave * R + sub, nwaves * R, n, j, sub, Lptr, Lidx, Lval,
                        [&](int i) { return x[(int64_t)i * K + j]; }, [&](int row, double acc) {
                          if (run) {
                            const double t = r0[(int64_t)row * K + j] - acc;
                            r[(int64_t)row * K + j] = t;
                            a[0] += t * t;
                          }
                        });
  const int slots[1] = {S_RR};
  mblock_sum<K, 1>(a, part, slots);
}

// After the solve: (b - bmu) - la x in FP64 for each returned column
// (x = the iterate k_mfinal returned: x when its copy is pending, else
// bestx; b from the caller's column-major B through perm), partial sums of
// squares per column
template <int K>
__global__ void __launch_bounds__(kBlock)
k_mtrue_resid(int n, int k, const int* __restrict__ perm, const int* __restrict__ Lptr,
              const int* __restrict__ Lidx, const double* __restrict__ Lval,
              const double* __restrict__ x, const double* __restrict__ bestx,
              const double* __restrict__ B, double* __restrict__ part,
              const DevState* __restrict__ mst) {
  constexpr int R = kWave / K;
  const int lane = threadIdx.x % kWave, j = lane % K, sub = lane / K;
  const bool run = j < k;
  const double* __restrict__ src = mst[j].pending_copy ? x : bestx;
  const double mu = mst[j].bmu;
  const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int nwaves = gridDim.x * (kBlock / kWave);
  double a[1] = {0.0};
  mrows_in_flight<K, 2>(wave * R + sub, nwaves * R, n, j, sub, Lptr, Lidx, Lval,
                        [&](int i) { return src[(int64_t)i * K + j]; }, [&](int row, double acc) {
                          if (run) {
                            const double t = (B[(int64_t)j * n + perm[row]] - mu) - acc;
                            a[0] += t * t;
                          }
                        });
  const int slots[1] = {S_RR};
  mblock_sum<K, 1>(a, part, slots);
}

template <int K>
__global__ void k_mscalar_true(const double* __restrict__ part, DevState* mst) {
  const double rr = mtotal<K>(part, S_RR);
  if (threadIdx.x < K) {
    DevState* st = mst + threadIdx.x;
    st->true_nr = st->nb > 0.0 ? sqrt(rr) / st->nb : 0.0;
  }
}

// after k_mresid: the true relative residual replaces nr (as
// k_scalar_resid, per column); convergence is accepted only on it
template <int K>
__global__ void k_mscalar_resid(int always, const double* __restrict__ part, DevState* mst,
                                int* flags) {
  if (flags[0] || !(always || flags[3])) return;
  const double rr = mtotal<K>(part, S_RR);
  if (threadIdx.x < K) {
    DevState* st = mst + threadIdx.x;
    if (!st->stopped && (always || st->verify)) {
      const double nr = sqrt(rr) / st->nb;
      st->nr = nr;
      if (!st->spmv64 && nr < st->switch_tol) st->spmv64 = 1;
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
  mark_all<K>(mst, flags);
  if (threadIdx.x == 0) flags[3] = 0;
}

template <int K>
__global__ void k_mscalar_rho(const double* __restrict__ part, DevState* mst, int* flags) {
  if (flags[0]) return;
  const double zr = mtotal<K>(part, S_ZR);
  const double sz = mtotal<K>(part, S_AZ);
  const double sp = mtotal<K>(part, S_AP2);
  if (threadIdx.x < K && !mst[threadIdx.x].stopped) step_rho(mst + threadIdx.x, zr, sz, sp);
  mark_all<K>(mst, flags);
}

// p = (yF - mu) + beta p for the right-hand sides still running (and its
// FP32 copy when pf)
template <int K, typename T>
__global__ void __launch_bounds__(kBlock)
k_mpupdate(int n, const T* __restrict__ z, double* __restrict__ p, float* __restrict__ pf,
           const DevState* __restrict__ mst, const int* __restrict__ flags) {
  if (flags[0]) return;
  const int j = threadIdx.x % K;
  if (mst[j].stopped) return;
  const double mu = mst[j].mu, beta = mst[j].beta;
  const int64_t nK = (int64_t)n * K;
  const int64_t S = (int64_t)gridDim.x * kBlock;
  int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  for (; e + 3 * S < nK; e += 4 * S) {
    double zv[4], pv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      zv[u] = (double)z[e + u * S];
      pv[u] = p[e + u * S];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double t = (zv[u] - mu) + beta * pv[u];
      p[e + u * S] = t;
      if (pf) pf[e + u * S] = (float)t;
    }
  }
  for (; e < nK; e += S) {
    const double t = ((double)z[e] - mu) + beta * p[e];
    p[e] = t;
    if (pf) pf[e] = (float)t;
  }
}

// X[j*n + perm[v]] = best iterate (x where the last iteration improved)
template <int K>
__global__ void __launch_bounds__(kBlock)
k_mfinal(int n, int k, const int* __restrict__ perm, const double* __restrict__ x,
         const double* __restrict__ bestx, double* __restrict__ X,
         const DevState* __restrict__ mst) {
  const int j = threadIdx.x % K;
  if (j >= k) return;
  const bool cp = mst[j].pending_copy != 0;
  const double* src = cp ? x : bestx;
  const int64_t nK = (int64_t)n * K;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nK;
       e += (int64_t)gridDim.x * kBlock)
    X[(int64_t)j * n + perm[(int)(e / K)]] = src[e];
}

}  // namespace lapgpu
