// lapgpu_wg.hpp — one whole PCG solve per workgroup (gfx950).
//
// For batches of small graphs and many right-hand sides: workgroup i runs
// the complete solve closure of src/approxChol.jl:1624 (pcg, src/pcg.jl:
// 256-445, with LDLsolver, src/approxChol.jl:1347-1408) for request i, with
// no grid-wide synchronisation and no host round trip.  The LDLinv sweeps of
// the whole graph are one slot program (see lapgpu_kernels.hpp); y lives in
// LDS when the graph has at most kTailCapV vertices, in global memory
// otherwise.  Every scalar decision is computed redundantly by all threads
// from block-reduced values, so control flow stays uniform.
#pragma once

#include "lapgpu_device.hpp"

namespace lapgpu {

constexpr int kWgThreads = 1024;
constexpr int kWgG = 8;  // lanes per SpMV row
// entries per lane of the whole-graph slot programs (their max_elog is
// kWgElog): one, so the solver keeps its registers for the rest
constexpr int kWgElog = 0;
constexpr int kWgPlanes = 1 << kWgElog;

struct WgGraphD {
  int n, nf, nb, npart;  // npart: partial slots of the programs, past y[n)
  const int* perm;
  const int* Lptr;
  const int* Lidx;
  const double* Lval;
  const double* s;
  const TailStepD* fsteps;
  const TailSlotD* fslots;
  const TailStepD* bsteps;
  const TailSlotD* bslots;
};

struct WgReqD {
  const double* b;  // rhs, caller's vertex order
  double* x;        // solution, caller's vertex order
  double* work;     // 8 n + npart doubles
  double tol;
  long long maxits;
  unsigned long long maxticks;
  int graph, stag_test, center, has_maxtime;
  // cancel flag in coherent host memory (nullptr: not cancellable); read
  // once per iteration
  const int* cancel;
  // results
  long long its;
  double nr;
  int stop, pad;
};

// Sum NV values over the workgroup; every thread receives the totals.
template <int NV>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double* red) {
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  constexpr int W = kWgThreads / kWave;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = group_sum<kWave>(v[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) red[i * W + wid] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) s += red[i * W + w];
    v[i] = s;
  }
  __syncthreads();
}

// z = LDLsolver(r) (without mean removal: y is returned in ybuf); the mean
// removal is fused into the caller's next pass.
template <bool LDSY, typename Y>
__device__ __forceinline__ void wg_ldl(const WgGraphD& G, const double* r, Y* y) {
  const int n = G.n;
  for (int i = threadIdx.x; i < n; i += kWgThreads) y[i] = r[i];
  __syncthreads();
  run_program<4, LDSY, kWgPlanes>(G.fsteps, G.nf, G.fslots, y);
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kWgThreads) y[i] *= G.s[i];
  __syncthreads();
  run_program<4, LDSY, kWgPlanes>(G.bsteps, G.nb, G.bslots, y);
  __syncthreads();
}

template <bool LDSY, typename Y>
__device__ __forceinline__ void wg_pcg(const WgGraphD& G, WgReqD& R, Y* y, double* red) {
  const int n = G.n;
  const int t = threadIdx.x;
  double* bp = R.work;
  double* r = bp + n;
  double* x = r + n;
  double* bestx = x + n;
  double* p = bestx + n;
  double* q = p + n;
  double* z = q + n;

  // b .- mean(b) in level order (src/approxChol.jl:1624)
  {
    double v[1] = {0.0};
    for (int i = t; i < n; i += kWgThreads) {
      const double bi = R.b[G.perm[i]];
      bp[i] = bi;
      v[0] += bi;
    }
    wg_sum<1>(v, red);
    const double mu = R.center ? v[0] / (double)n : 0.0;
    double w[1] = {0.0};
    for (int i = t; i < n; i += kWgThreads) {
      const double ri = bp[i] - mu;
      r[i] = ri;
      x[i] = 0.0;
      bestx[i] = 0.0;
      w[0] += ri * ri;
    }
    wg_sum<1>(w, red);
    R.nr = 1.0;
    const double nb = sqrt(w[0]);
    if (nb == 0.0) {
      for (int i = t; i < n; i += kWgThreads) R.x[G.perm[i]] = 0.0;
      R.its = 0;
      R.stop = 9;
      return;
    }
    // z = pre(r); p = z; rho = r.z
    wg_ldl<LDSY>(G, r, y);
    double m[1] = {0.0};
    for (int i = t; i < n; i += kWgThreads) m[0] += y[i];
    wg_sum<1>(m, red);
    const double my = m[0] / (double)n;
    double zr[1] = {0.0};
    for (int i = t; i < n; i += kWgThreads) {
      const double zi = y[i] - my;
      p[i] = zi;
      zr[0] += r[i] * zi;
    }
    wg_sum<1>(zr, red);

    double rho = zr[0], best_rho = rho, bestnr = 1.0, nr = 1.0;
    const double stag_fac = 1.0 - 1.0 / (double)R.stag_test;
    long long itcnt = 0, stag_count = 0;
    int stop = 1;
    bool pending = false;
    const unsigned long long t1 = wall_clock64();
    const bool poll = R.has_maxtime || R.cancel != nullptr;
    while (R.maxits < 0 || itcnt < R.maxits) {
      itcnt += 1;
      const bool chk = (itcnt % 10) == 0;
      // the cancel flag, read at the start of the iteration (the read's
      // latency hides behind the SpMV) and acted on at its end
      int cflag = 0;
      if (R.cancel != nullptr && t == 0)
        cflag = __hip_atomic_load(R.cancel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      // q = L p, p.q
      double pq[1] = {0.0};
      {
        const int grp = t / kWgG, lane = t % kWgG;
        rows_in_flight<kWgG, 2>(grp, kWgThreads / kWgG, n, G.Lptr, G.Lidx, G.Lval,
                                [&](int i) { return p[i]; }, [&](int row, double acc) {
                                  if (lane == 0) {
                                    q[row] = acc;
                                    pq[0] += p[row] * acc;
                                  }
                                });
      }
      wg_sum<1>(pq, red);
      const double al = rho / pq[0];
      if (chk) {
        double a[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = t; i < n; i += kWgThreads) {
          a[0] += fabs(p[i]);
          a[1] += fabs(x[i]);
          a[2] += fabs(q[i]);
          a[3] += fabs(r[i]);
        }
        wg_sum<4>(a, red);
        if (al * a[0] < kEps * a[1]) {
          stop = 4;
          break;
        }
        if (al * a[2] < kEps * a[3]) {
          // the reference updates x before this test; bestx only changes
          // after it, so stopping here returns the same vector
          stop = 5;
          break;
        }
      }
      double rr[1] = {0.0};
      for (int i = t; i < n; i += kWgThreads) {
        const double xo = x[i];
        if (pending) bestx[i] = xo;
        x[i] = xo + al * p[i];
        const double rn = r[i] + (-al) * q[i];
        r[i] = rn;
        rr[0] += rn * rn;
      }
      wg_sum<1>(rr, red);
      pending = false;
      nr = sqrt(rr[0]) / nb;
      if (nr < bestnr) {
        bestnr = nr;
        pending = true;
      }
      if (nr < R.tol) {
        stop = 0;
        break;
      }
      wg_ldl<LDSY>(G, r, y);
      double m2[1] = {0.0};
      for (int i = t; i < n; i += kWgThreads) m2[0] += y[i];
      wg_sum<1>(m2, red);
      const double my2 = m2[0] / (double)n;
      double zv[3] = {0.0, 0.0, 0.0};
      for (int i = t; i < n; i += kWgThreads) {
        const double zi = y[i] - my2;
        z[i] = zi;
        zv[0] += zi * r[i];
        if (chk) {
          zv[1] += fabs(zi);
          zv[2] += fabs(p[i]);
        }
      }
      wg_sum<3>(zv, red);
      const double oldrho = rho;
      rho = zv[0];
      if (rho < best_rho * stag_fac) {
        best_rho = rho;
        stag_count = 0;
      } else if (R.stag_test > 0 && best_rho > stag_fac * rho) {
        stag_count += 1;
        if (stag_count > R.stag_test) {
          stop = 3;
          break;
        }
      }
      if (isinf(rho)) {
        stop = 7;
        break;
      }
      const double beta = rho / oldrho;
      if (isinf(beta)) {
        stop = 8;
        break;
      }
      if (chk && beta * zv[2] < kEps * zv[1]) {
        stop = 6;
        break;
      }
      for (int i = t; i < n; i += kWgThreads) p[i] = z[i] + beta * p[i];
      if (poll) {
        // one clock reading and one flag, broadcast, so that every thread
        // agrees
        if (t == 0)
          red[0] = (R.has_maxtime && (wall_clock64() - t1) > R.maxticks) ? 1.0
                   : cflag != 0                                          ? 2.0
                                                                         : 0.0;
        __syncthreads();
        const double why = red[0];
        __syncthreads();
        if (why == 1.0) {
          stop = 2;
          break;
        }
        if (why == 2.0) {
          stop = 10;
          break;
        }
      } else {
        __syncthreads();
      }
    }
    if (pending) {
      for (int i = t; i < n; i += kWgThreads) bestx[i] = x[i];
    }
    __syncthreads();
    for (int i = t; i < n; i += kWgThreads) R.x[G.perm[i]] = bestx[i];
    R.its = itcnt;
    R.nr = nr;
    R.stop = stop;
  }
}

__global__ void __launch_bounds__(kWgThreads)
k_wg_pcg(const WgGraphD* __restrict__ graphs, WgReqD* __restrict__ reqs) {
  __shared__ double ly[kTailCapV + kPartSlots];
  __shared__ double red[4 * (kWgThreads / kWave)];
  WgReqD R = reqs[blockIdx.x];
  const WgGraphD G = graphs[R.graph];
  if (G.n <= kTailCapV)
    wg_pcg<true>(G, R, ly, red);
  else
    wg_pcg<false>(G, R, R.work + 7 * (size_t)G.n, red);
  if (threadIdx.x == 0) {
    reqs[blockIdx.x].its = R.its;
    reqs[blockIdx.x].nr = R.nr;
    reqs[blockIdx.x].stop = R.stop;
  }
}

}  // namespace lapgpu
