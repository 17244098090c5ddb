/*
 * lap_oracle.c — CPU restatement of the reference's approxchol_lap SOLVE
 * phase.  TEST INFRASTRUCTURE ONLY: the parity checker for liblapgpu.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it; the product never links or calls it.
 *
 * Each function follows the reference line by line, in the same arithmetic
 * order, with 0-based indices:
 *   orc_spmv_csc      SparseArrays CSC `mat*p` as called at src/pcg.jl:291
 *   orc_forward       forward!   src/approxChol.jl:1369-1388
 *   orc_backward      backward!  src/approxChol.jl:1390-1408
 *   orc_ldlsolver     LDLsolver  src/approxChol.jl:1347-1366
 *   orc_pcg           pcg        src/pcg.jl:256-445 (+ axpy2!/bzbeta! :448-461)
 *   orc_approxchol_solve  the solve closure src/approxChol.jl:1624
 *
 * Pinning (tests/test_oracle.py): SpMV against scipy and dense products;
 * LDLsolver against the Moore-Penrose pseudo-inverse on trees and cycles,
 * where the elimination is exact; forward/backward against the closed form
 * of the sweeps and each other (backward is forward's transpose); pcg
 * against a textbook PCG (iterates and iteration counts), an exact
 * preconditioner (one iteration), b = 0, maxits = 1 and maxtime.
 * BLAS dot/nrm2 and Statistics.mean are restated as plain left-to-right
 * sums (the reference's OpenBLAS kernels are not pinned by any test).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_EPS 2.220446049250313e-16 /* eps(Float64) */

enum {
  ORC_CONVERGED = 0,
  ORC_MAXITS = 1,
  ORC_MAXTIME = 2,
  ORC_STAG_TEST = 3,
  ORC_STAG_AL_P = 4,
  ORC_STAG_AL_Q = 5,
  ORC_STAG_BETA = 6,
  ORC_INF_RHO = 7,
  ORC_INF_BETA = 8,
  ORC_ZERO_RHS = 9
};

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void orc_spmv_csc(int64_t n, const int64_t* colptr, const int64_t* rowval,
                  const double* nzval, const double* x, double* y) {
  for (int64_t i = 0; i < n; ++i) y[i] = 0.0;
  for (int64_t j = 0; j < n; ++j) {
    const double xj = x[j];
    for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k)
      y[rowval[k]] += nzval[k] * xj;
  }
}

void orc_forward(int64_t n, const int64_t* col, const int64_t* colptr,
                 const int64_t* rowval, const double* fval, double* y) {
  for (int64_t ii = 0; ii < n - 1; ++ii) {
    const int64_t i = col[ii];
    const int64_t j0 = colptr[ii];
    const int64_t j1 = colptr[ii + 1] - 1;
    double yi = y[i];
    for (int64_t jj = j0; jj < j1; ++jj) {
      const int64_t j = rowval[jj];
      y[j] += fval[jj] * yi;
      yi *= (1.0 - fval[jj]);
    }
    y[rowval[j1]] += yi;
    y[i] = yi;
  }
}

void orc_backward(int64_t n, const int64_t* col, const int64_t* colptr,
                  const int64_t* rowval, const double* fval, double* y) {
  for (int64_t ii = n - 2; ii >= 0; --ii) {
    const int64_t i = col[ii];
    const int64_t j0 = colptr[ii];
    const int64_t j1 = colptr[ii + 1] - 1;
    double yi = y[i];
    yi = yi + y[rowval[j1]];
    for (int64_t jj = j1 - 1; jj >= j0; --jj) {
      const int64_t j = rowval[jj];
      yi = (1.0 - fval[jj]) * yi + fval[jj] * y[j];
    }
    y[i] = yi;
  }
}

void orc_ldlsolver(int64_t n, const int64_t* col, const int64_t* colptr,
                   const int64_t* rowval, const double* fval, const double* d,
                   const double* b, double* y) {
  memcpy(y, b, sizeof(double) * (size_t)n);
  orc_forward(n, col, colptr, rowval, fval, y);
  for (int64_t i = 0; i < n; ++i)
    if (d[i] != 0.0) y[i] /= d[i];
  orc_backward(n, col, colptr, rowval, fval, y);
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += y[i];
  const double mu = s / (double)n;
  for (int64_t i = 0; i < n; ++i) y[i] = y[i] - mu;
}

/* ---- pcg ----------------------------------------------------------------- */

typedef void (*orc_pre_fn)(void* ctx, int64_t n, const double* r, double* z);

typedef struct {
  const int64_t *col, *colptr, *rowval;
  const double *fval, *d;
} orc_ldl_ctx;

static void pre_ldl(void* c, int64_t n, const double* r, double* z) {
  const orc_ldl_ctx* L = (const orc_ldl_ctx*)c;
  orc_ldlsolver(n, L->col, L->colptr, L->rowval, L->fval, L->d, r, z);
}

static void pre_diag(void* c, int64_t n, const double* r, double* z) {
  const double* dg = (const double*)c;
  for (int64_t i = 0; i < n; ++i) z[i] = r[i] / dg[i];
}

static double dot(int64_t n, const double* a, const double* b) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static double asum(int64_t n, const double* a) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += fabs(a[i]);
  return s;
}

/* Result record of one pcg call. */
typedef struct {
  int64_t its;      /* itcnt, written to pcgIts[1] by the reference */
  int32_t stop;     /* ORC_* reason */
  int32_t pad;
  double nr;        /* last recursive relative residual norm(r)/nb */
  double bestnr;
  double rho;
} orc_pcg_result;

static int orc_pcg_core(int64_t n, const int64_t* colptr, const int64_t* rowval,
                        const double* nzval, const double* b, orc_pre_fn pre,
                        void* pctx, double tol, int64_t maxits, double maxtime,
                        int64_t stag_test, double* bestx, orc_pcg_result* res) {
  memset(res, 0, sizeof(*res));
  double nb = sqrt(dot(n, b, b));
  if (nb == 0.0) {
    memset(bestx, 0, sizeof(double) * (size_t)n);
    res->stop = ORC_ZERO_RHS;
    return 0;
  }
  double* x = (double*)calloc((size_t)n, sizeof(double));
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  if (!x || !r || !z || !p || !q) {
    free(x); free(r); free(z); free(p); free(q);
    return -1;
  }
  memset(bestx, 0, sizeof(double) * (size_t)n);
  double bestnr = 1.0;
  memcpy(r, b, sizeof(double) * (size_t)n);
  pre(pctx, n, r, z);
  memcpy(p, z, sizeof(double) * (size_t)n);
  double rho = dot(n, r, z);
  double best_rho = rho;
  int64_t stag_count = 0;
  /* Julia: 1 - 1/stag_test with stag_test an Integer is a Float64; for
   * stag_test == 0 it is -Inf. */
  const double stag_fac = 1.0 - 1.0 / (double)stag_test;
  const double t1 = now_s();
  int64_t itcnt = 0;
  int stop = ORC_MAXITS;
  double nr = 1.0;

  while (itcnt < maxits) {
    itcnt += 1;
    orc_spmv_csc(n, colptr, rowval, nzval, p, q);
    const double pq = dot(n, p, q);
    const double al = rho / pq;
    if (itcnt % 10 == 0) {
      if (al * asum(n, p) < ORC_EPS * asum(n, x)) { stop = ORC_STAG_AL_P; break; }
    }
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + al * p[i];
    if (itcnt % 10 == 0) {
      if (al * asum(n, q) < ORC_EPS * asum(n, r)) { stop = ORC_STAG_AL_Q; break; }
    }
    for (int64_t i = 0; i < n; ++i) r[i] = r[i] + (-al) * q[i];
    nr = sqrt(dot(n, r, r)) / nb;
    if (nr < bestnr) {
      bestnr = nr;
      memcpy(bestx, x, sizeof(double) * (size_t)n);
    }
    if (nr < tol) { stop = ORC_CONVERGED; break; }

    pre(pctx, n, r, z);
    const double oldrho = rho;
    rho = dot(n, z, r);
    if (rho < best_rho * stag_fac) {
      best_rho = rho;
      stag_count = 0;
    } else if (stag_test > 0) {
      if (best_rho > stag_fac * rho) {
        stag_count += 1;
        if (stag_count > stag_test) { stop = ORC_STAG_TEST; break; }
      }
    }
    if (isinf(rho)) { stop = ORC_INF_RHO; break; }
    const double beta = rho / oldrho;
    if (isinf(beta)) { stop = ORC_INF_BETA; break; }
    if (itcnt % 10 == 0) {
      if (beta * asum(n, p) < ORC_EPS * asum(n, z)) { stop = ORC_STAG_BETA; break; }
    }
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    if ((now_s() - t1) > maxtime) { stop = ORC_MAXTIME; break; }
  }
  res->its = itcnt;
  res->stop = stop;
  res->nr = nr;
  res->bestnr = bestnr;
  res->rho = rho;
  free(x); free(r); free(z); free(p); free(q);
  return 0;
}

/* pcg(la, b, LDLsolver) — the preconditioned solve without the closure's
 * mean removal. */
int orc_pcg_ldl(int64_t n, const int64_t* colptr, const int64_t* rowval,
                const double* nzval, const int64_t* col, const int64_t* fcolptr,
                const int64_t* frowval, const double* fval, const double* d,
                const double* b, double tol, int64_t maxits, double maxtime,
                int64_t stag_test, double* x, orc_pcg_result* res) {
  orc_ldl_ctx c = {col, fcolptr, frowval, fval, d};
  return orc_pcg_core(n, colptr, rowval, nzval, b, pre_ldl, &c, tol, maxits,
                      maxtime, stag_test, x, res);
}

/* pcg(la, b, x -> x ./ dg) — the Jacobi setup of test/testPCG.jl:1-11. */
int orc_pcg_diag(int64_t n, const int64_t* colptr, const int64_t* rowval,
                 const double* nzval, const double* dg, const double* b,
                 double tol, int64_t maxits, double maxtime, int64_t stag_test,
                 double* x, orc_pcg_result* res) {
  return orc_pcg_core(n, colptr, rowval, nzval, b, pre_diag, (void*)dg, tol,
                      maxits, maxtime, stag_test, x, res);
}

/* f(b) = pcg(la, b .- mean(b), F; stag_test) — src/approxChol.jl:1624. */
int orc_approxchol_solve(int64_t n, const int64_t* colptr,
                         const int64_t* rowval, const double* nzval,
                         const int64_t* col, const int64_t* fcolptr,
                         const int64_t* frowval, const double* fval,
                         const double* d, const double* b, double tol,
                         int64_t maxits, double maxtime, int64_t stag_test,
                         double* x, orc_pcg_result* res) {
  double* bc = (double*)malloc(sizeof(double) * (size_t)n);
  if (!bc) return -1;
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += b[i];
  const double mu = s / (double)n;
  for (int64_t i = 0; i < n; ++i) bc[i] = b[i] - mu;
  int rc = orc_pcg_ldl(n, colptr, rowval, nzval, col, fcolptr, frowval, fval,
                       d, bc, tol, maxits, maxtime, stag_test, x, res);
  free(bc);
  return rc;
}
