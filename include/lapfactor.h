/*
 * lapfactor.h — host-side approximate-Cholesky factor builder (LDLinv).
 *
 * In a deployment with Julia on the host, the factor is built by the
 * reference's own `approxChol(LLmatp(a))` (src/approxChol.jl:864-977) and
 * handed to liblapgpu through `lapgpu_upload_ldli`.  Where no Julia is present
 * (this image, the serving engine's "new graph" admissions), this library
 * restates that elimination in C++ with its own documented RNG stream
 * (xoshiro256++ seeded through splitmix64; uniform doubles = top 53 bits).
 * The output has exactly the reference's LDLinv layout
 * (src/approxCholTypes.jl:73-79): col (n-1), colptr (n), rowval/fval (nnzF),
 * d (n, indexed by vertex; the never-eliminated vertex has d = 0).
 *
 * The library is re-entrant: independent builds may run on separate threads.
 */
#ifndef LAPFACTOR_H
#define LAPFACTOR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lapfactor_ldli lapfactor_ldli;

/* Build an LDLinv from a symmetric, non-negative adjacency matrix with zero
 * diagonal in CSC form (rows sorted inside each column is not required).
 * `index_base` is 0 or 1 and applies to colptr and rowval.
 * Mirrors `approxChol(LLmatp(a))` with ApproxCholParams(:deg, 5, 0, 0)
 * (src/approxChol.jl:87, 102-136, 864-977).  Returns 0 on success; a
 * negative code on invalid input (negative weight: the check of
 * src/approxChol.jl:1532-1534; isolated vertex; malformed CSC). */
int lapfactor_build(int64_t n, const int64_t* colptr, const int64_t* rowval,
                    const double* nzval, int index_base, uint64_t seed,
                    lapfactor_ldli** out);

/* The fixed-order elimination approxChol(LLMatOrd(a, p))
 * (src/approxChol.jl:203-231, 265-296, 758-862), which ApproxCholParams
 * orders :given (p = identity), :wdeg (p = sortperm of weighted degree times
 * 1 + U(0,1)) and :random (a random permutation) use
 * (src/approxChol.jl:1628-1790).  perm[t] is the vertex eliminated t-th
 * (NULL: the given order 0..n-1); `index_base` applies to perm too.  The
 * factor is returned in the graph's own vertex labels (col, rowval and d
 * relabelled as the reference's ldlip), so it uploads like any other.  Same
 * RNG stream and error codes as lapfactor_build. */
int lapfactor_build_ordered(int64_t n, const int64_t* colptr, const int64_t* rowval,
                            const double* nzval, int index_base, const int64_t* perm,
                            uint64_t seed, lapfactor_ldli** out);

/* The split-edge elimination approxChol(LLmatp(a, split), split, merge)
 * (src/approxChol.jl:138-186, 541-603, 979-1157, 2121-2144) that
 * approxchol_lapGreedy runs when ApproxCholParams has split >= 1 and
 * merge >= 1, and approxchol_lap2 requests (src/approxChol.jl:1559-1579).
 * Every edge starts as `split` copies of weight w/split; a column's copies
 * of one row share their mean weight, at most `merge` of them are kept
 * (merge < 1: all of them, the split-only elimination of
 * src/approxChol.jl:1162-1339), and the factor holds one entry per row.
 * Same RNG stream and error codes as lapfactor_build. */
int lapfactor_build_split(int64_t n, const int64_t* colptr, const int64_t* rowval,
                          const double* nzval, int index_base, int64_t split, int64_t merge,
                          uint64_t seed, lapfactor_ldli** out);

/* Number of vertices n and of factor entries nnzF = len(fval). */
int64_t lapfactor_n(const lapfactor_ldli* f);
int64_t lapfactor_nnz(const lapfactor_ldli* f);

/* Copy the factor out.  Lengths: col n-1, colptr n, rowval nnzF, fval nnzF,
 * d n.  `index_base` (0 or 1) applies to col, colptr and rowval. */
int lapfactor_export(const lapfactor_ldli* f, int index_base, int64_t* col,
                     int64_t* colptr, int64_t* rowval, double* fval,
                     double* d);

void lapfactor_free(lapfactor_ldli* f);

/* Message of the last failure on the calling thread ("" if none). */
const char* lapfactor_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LAPFACTOR_H */
