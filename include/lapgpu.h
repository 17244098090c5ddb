/*
 * lapgpu.h — C-ABI of the MI355X (gfx950) solve phase of approxchol_lap.
 *
 * The reference's solve is the closure returned by approxchol_lapGreedy
 * (src/approxChol.jl:1624):
 *     f(b; tol, maxits, maxtime, verbose, pcgIts) =
 *         pcg(la, b .- mean(b), F; tol, maxits, maxtime, verbose, pcgIts,
 *             stag_test = params.stag_test)
 * with F(b) = LDLsolver(ldli, b) (src/approxChol.jl:1347-1408, 1614) and pcg
 * from src/pcg.jl:256-445.  A host (Julia via ccall, Python via ctypes, C++)
 * builds `la = lap(a)` and the LDLinv factor, uploads both once, then calls
 * lapgpu_solve per right-hand side.  Everything inside lapgpu_solve runs on
 * the GPU: the Laplacian SpMV, the LDLinv forward/backward sweeps and PCG's
 * reductions, updates and stopping rules.
 *
 * Conventions
 *  - Return codes: 0 = success; < 0 = failure, with a message available from
 *    lapgpu_last_error(ctx) (or lapgpu_last_error(NULL) when ctx creation
 *    failed).  The Julia shim turns a nonzero code into `error(...)`, the
 *    reference's error style (src/approxChol.jl:1533).
 *  - Index arrays are int64 with an explicit base (1 for Julia's native
 *    arrays, 0 for C/Python); they are converted on upload.
 *  - All host buffers belong to the caller; uploads deep-copy.
 *  - One host thread per context at a time.
 */
#ifndef LAPGPU_H
#define LAPGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lapgpu_ctx lapgpu_ctx;

/* Stop reasons, one per exit of pcg (src/pcg.jl:269-430). */
enum lapgpu_stop {
  LAPGPU_CONVERGED = 0,   /* nr < tol                          pcg.jl:355  */
  LAPGPU_MAXITS = 1,      /* itcnt == maxits                   pcg.jl:288  */
  LAPGPU_MAXTIME = 2,     /* time() - t1 > maxtime             pcg.jl:425  */
  LAPGPU_STAG_TEST = 3,   /* stag_count > stag_test            pcg.jl:370  */
  LAPGPU_STAG_AL_P = 4,   /* al*sum|p| < eps*sum|x|            pcg.jl:308  */
  LAPGPU_STAG_AL_Q = 5,   /* al*sum|q| < eps*sum|r|            pcg.jl:330  */
  LAPGPU_STAG_BETA = 6,   /* beta*sum|p| < eps*sum|z|          pcg.jl:406  */
  LAPGPU_INF_RHO = 7,     /* isinf(rho)                        pcg.jl:385  */
  LAPGPU_INF_BETA = 8,    /* isinf(beta)                       pcg.jl:397  */
  LAPGPU_ZERO_RHS = 9,    /* nb == 0: returns zeros            pcg.jl:269  */
  LAPGPU_CANCELLED = 10   /* lapgpu_cancel / lapgpu_batch_cancel: stopped at
                             an iteration boundary, best iterate returned (no
                             reference counterpart; the reference's only
                             interruption is its time limit, maxtime) */
};

/* lapgpu_solve flags */
#define LAPGPU_VERBOSE 1      /* print the reference's verbose lines        */
#define LAPGPU_NO_CENTER 2    /* skip the closure's b .- mean(b) (plain pcg) */
#define LAPGPU_EAGER 8        /* launch this solve's iterations directly; do
                                 not capture them as a HIP graph (a context
                                 solved once skips the capture; no reference
                                 counterpart)                                */

/* Precision modes (lapgpu_set_precision). */
#define LAPGPU_FP64 0         /* everything in FP64 (the reference's path)   */
#define LAPGPU_FP32_INNER 1   /* SpMV + LDLinv apply in FP32; dots, alpha,
                                 beta and a true-residual replacement every
                                 `replace_every` iterations in FP64          */

/* Create a context on HIP device `device`: one stream of default priority
 * and an empty graph slot.  No reference counterpart (the reference has no
 * device); this is the handle the Julia shim keeps in its closure. */
int lapgpu_create(lapgpu_ctx** out, int device);

/* Upload la = lap(a) (src/graphOps.jl:25), the operand of `mat*p`
 * (src/pcg.jl:291).  CSC of an n-by-n symmetric matrix; since it is
 * symmetric the device treats it as CSR.  Replaces SparseArrays' CSC SpMV. */
int lapgpu_upload_csc(lapgpu_ctx* ctx, int64_t n, const int64_t* colptr,
                      const int64_t* rowval, const double* nzval,
                      int index_base);

/* Upload an LDLinv (src/approxCholTypes.jl:73-79): col (n-1), colptr (n),
 * rowval/fval (nnzF), d (n, by vertex).  Computes on the host the level
 * schedule of the column DAG and the per-entry sweep weights, then copies
 * the level-ordered factor to the device.  Replaces the data that
 * `F(b) = LDLsolver(ldli, b)` (src/approxChol.jl:1614) closes over. */
int lapgpu_upload_ldli(lapgpu_ctx* ctx, int64_t n, int64_t nnzF,
                       const int64_t* col, const int64_t* colptr,
                       const int64_t* rowval, const double* fval,
                       const double* d, int index_base);

/* Solve la x = b with host buffers b (n) and x_out (n): the closure
 * f(b; tol, maxits, maxtime, verbose, pcgIts) of src/approxChol.jl:1624.
 * Returns pcg's bestx (src/pcg.jl:444).  its_out receives itcnt (what
 * pcgIts[1] receives, src/pcg.jl:439-441), relres_out the last recursive
 * relative residual norm(r)/nb, stop_reason_out a lapgpu_stop value.  Any
 * of the three out pointers may be NULL.  maxits < 0 means no limit and
 * maxits == 0 runs no iteration (returns zeros, as the reference does);
 * maxtime <= 0 or infinite means no limit. */
int lapgpu_solve(lapgpu_ctx* ctx, const double* b, double* x_out, int64_t n,
                 double tol, int64_t maxits, double maxtime, int stag_test,
                 int flags, int64_t* its_out, double* relres_out,
                 int* stop_reason_out);

/* As lapgpu_solve, with b and x_out in device memory (resident inputs). */
int lapgpu_solve_device(lapgpu_ctx* ctx, const double* d_b, double* d_x,
                        int64_t n, double tol, int64_t maxits, double maxtime,
                        int stag_test, int flags, int64_t* its_out,
                        double* relres_out, int* stop_reason_out);

/* k right-hand sides of the same graph at once: B and X_out are n-by-k,
 * column-major (column j at B + j*n, a Julia Matrix{Float64}).  Column j
 * gets the result of lapgpu_solve(B[:, j]) with its own stopping rules
 * (itcnt, relres and stop reason in its_out[j], relres_out[j],
 * stop_reason_out[j]; arrays of k, each may be NULL), all columns sharing
 * every launch.  The reference solves one vector per call of the closure
 * (src/approxChol.jl:1624); this is the same closure applied column by
 * column.  Columns are processed 64 at a time, in the context's precision
 * mode (lapgpu_set_precision), as lapgpu_solve. */
int lapgpu_solve_multi(lapgpu_ctx* ctx, const double* B, double* X_out, int64_t n,
                       int64_t k, double tol, int64_t maxits, double maxtime,
                       int stag_test, int flags, int64_t* its_out,
                       double* relres_out, int* stop_reason_out);

/* As lapgpu_solve_multi, with B and X_out in device memory. */
int lapgpu_solve_multi_device(lapgpu_ctx* ctx, const double* d_B, double* d_X,
                              int64_t n, int64_t k, double tol, int64_t maxits,
                              double maxtime, int stag_test, int flags,
                              int64_t* its_out, double* relres_out,
                              int* stop_reason_out);

/* The FP64 true relative residual ||(b - mean b) - la x|| / ||b - mean b||
 * (no mean removed under LAPGPU_NO_CENTER) of each column that the last
 * solve call on ctx returned: k = 1 after lapgpu_solve(_device), the call's
 * k after lapgpu_solve_multi(_device); -1 for another k.  It is computed on
 * the device after the solve (one more SpMV, in FP64 whatever the
 * precision mode), so in FP32-inner mode it is the per-solve error
 * against FP64 that the precision form reports (convergence there is
 * accepted only on this residual).  The reference's pcg reports only the
 * recursive residual (src/pcg.jl:434-437). */
int lapgpu_last_true_relres(lapgpu_ctx* ctx, double* out, int64_t k);

/* The plan file: everything lapgpu_upload_csc + lapgpu_upload_ldli build
 * on the host (level order, the permuted Laplacian, both sweep layouts, the
 * tail and dense-block slot programs, the level groups), so that a graph
 * seen before is admitted without planning again (the reference keeps the
 * factor as an LDLinv, src/approxCholTypes.jl:73-79, and saves solver
 * results with JLD2, compare/st_grid3.jl:71; it has no plan).
 * lapgpu_save_plan writes the context's plan (after its uploads or a
 * load) to `path` (written to path.tmp, then renamed).  lapgpu_load_plan
 * replaces both uploads: it reads a file written by lapgpu_save_plan,
 * refuses one whose checksum, version, shapes or indices do not fit (-1,
 * with a message), uploads it and writes the graph's n to *n_out (may be
 * NULL); the dense block's matrix is rebuilt on the device.  The file does not record the context settings it was
 * planned with: the loaded plan keeps them (group budget, dense block,
 * tail). */
int lapgpu_save_plan(lapgpu_ctx* ctx, const char* path);
int lapgpu_load_plan(lapgpu_ctx* ctx, const char* path, int64_t* n_out);

/* Cancel the solve running on ctx from another thread (no reference
 * counterpart; for serving, where a client may withdraw a request): column
 * j of the running lapgpu_solve_multi call, or every column when j == -1
 * (lapgpu_solve has one column, 0).  The solve stops those columns at the
 * next iteration boundary the host polls, with stop reason
 * LAPGPU_CANCELLED, their best iterate so far and its iteration count; the
 * other columns run on unchanged.  Returns 0 when a solve was running,
 * LAPGPU_IDLE when none was (nothing is cancelled), -1 for a column out of
 * range.  Thread-safe; the only entry point that may run concurrently with
 * a solve on the same context. */
#define LAPGPU_IDLE 2
int lapgpu_cancel(lapgpu_ctx* ctx, int64_t column);

/* Apply the preconditioner once: z = LDLsolver(ldli, r)
 * (src/approxChol.jl:1347-1366), host buffers. */
int lapgpu_apply_ldli(lapgpu_ctx* ctx, const double* r, double* z, int64_t n);

/* y = la * x once (src/pcg.jl:291), host buffers. */
int lapgpu_spmv(lapgpu_ctx* ctx, const double* x, double* y, int64_t n);

/* Select the precision mode for later solves (LAPGPU_FP64 by default). */
int lapgpu_set_precision(lapgpu_ctx* ctx, int mode, int replace_every);

/* How the calling thread waits for the device inside a solve (no reference
 * counterpart).  LAPGPU_WAIT_SPIN (default): HIP's polling wait, the lowest
 * latency, one busy core per waiting thread.  LAPGPU_WAIT_BLOCKING: the
 * thread queries the device's progress and sleeps 20 us between queries
 * (a wait ends up to ~20 us late), for processes that run many solves at
 * once on host threads and need the cores for other work (a serving
 * process factoring new graphs).  Call it between solves on the context. */
#define LAPGPU_WAIT_SPIN 0
#define LAPGPU_WAIT_BLOCKING 1
int lapgpu_set_wait_mode(lapgpu_ctx* ctx, int mode);

/* How narrow level groups of the single right-hand-side sweeps are launched
 * (no reference counterpart).  LAPGPU_SWEEP_CHAIN (default): each run of two
 * or more consecutive groups of at most 1024 rows (env LAPGPU_CHAIN_NARROW;
 * runs of fewer than 40 groups, env LAPGPU_CHAIN_MIN_RUN, stay one launch
 * per group, which measured faster for them) is one launch, whose workgroups take the groups' rows in sweep order and
 * wait for each gathered value to be written; LAPGPU_SWEEP_GROUPS: one
 * launch per group and sweep.  Both give the same bits.  Multi-RHS solves
 * always launch per group.  A chained run whose wait runs out
 * (it never should) switches the context to LAPGPU_SWEEP_GROUPS and the call
 * is run again; LAPGPU_CHAIN_RETRY is returned instead when its input is
 * gone (a device solve with d_x == d_b).  Env LAPGPU_CHAIN=0 sets
 * LAPGPU_SWEEP_GROUPS at context creation.  Call it between solves. */
#define LAPGPU_SWEEP_GROUPS 0
#define LAPGPU_SWEEP_CHAIN 1
#define LAPGPU_CHAIN_RETRY (-12)
int lapgpu_set_sweep_mode(lapgpu_ctx* ctx, int mode);

/* The chained runs of the uploaded schedule: how many runs, how many level
 * groups they hold (each sweep), and whether solves use them (active: 1;
 * 0 in LAPGPU_SWEEP_GROUPS mode or after a wait ran out).  No reference
 * counterpart.  Env LAPGPU_CHAIN_POLL_MAX (default 262144) bounds the polls
 * of one wait, for tests of the fallback. */
int lapgpu_chain_info(lapgpu_ctx* ctx, int64_t* runs, int64_t* chained_groups, int* active);

/* Level-group budget of the LDLinv schedule (no reference counterpart; the
 * reference has no device schedule).  The levels before the tail are run as
 * groups, one launch per group and sweep; a group spans at most max_depth
 * levels, its composite rows hold at most max_growth times the entries of
 * the rows they replace, and each level joined to a group adds at most
 * max_extra entries.  Defaults 32 / 16 / 196608 (measured on 3-D grids:
 * ~4 us per launch against ~8 ps per entry).  A smaller budget plans faster
 * on the host (fewer composite entries) at more launches per iteration:
 * 1 / 1 / 0 gives one group per level.  Applies to the next
 * lapgpu_upload_ldli; with a factor already uploaded the groups are rebuilt
 * at once (waits for the context's stream). */
int lapgpu_set_group_budget(lapgpu_ctx* ctx, int max_depth, double max_growth,
                            int64_t max_extra);

/* The dense tail block (no counterpart in the reference; part of this
 * library's LDLinv apply).  The deepest levels of the single-workgroup tail
 * are applied as one symmetric matrix D = (I - W^T)^-1 S (I - W)^-1 of their
 * K vertices (built on the device at upload by running their sweeps on unit
 * vectors), a matrix-vector product (many right-hand sides: a matrix-core
 * product) instead of their dependent steps.  The planner chooses K by a
 * cost model, at most max_k (0 disables; default 18432, the tail's LDS
 * capacity; env LAPGPU_DENSE_MAX).  Takes effect at the next
 * lapgpu_upload_ldli.  D costs 8 K^2 bytes (12 K^2 in the FP32-inner mode). */
int lapgpu_set_dense_tail(lapgpu_ctx* ctx, int64_t max_k);
/* K of the uploaded factor's dense block (0: none), the tail vertices
 * before it, and the device bytes of D. */
int lapgpu_dense_info(lapgpu_ctx* ctx, int64_t* k, int64_t* sparse_vertices, int64_t* bytes);

/* Factor entries (LDLinv nonzeros) each part of the schedule applies: the
 * forward and backward rows before the tail (the level groups), those of the
 * tail, and K of the dense block (measurement only). */
int lapgpu_sweep_entries(lapgpu_ctx* ctx, int64_t* groups_f, int64_t* groups_b, int64_t* tail_f,
                         int64_t* tail_b, int64_t* dense_k);

/* Schedule facts of the uploaded factor: number of levels, number of levels
 * run one launch each, the first vertex position of the single-workgroup
 * tail, nnzF. Any pointer may be NULL. */
int lapgpu_factor_info(lapgpu_ctx* ctx, int64_t* nlevels, int64_t* nwide,
                       int64_t* tail_start, int64_t* nnzF);

/* Per-iteration device times (ms) of the last LAPGPU_PROFILE solve's tail
 * with a dense block: the sparse forward, the dense product, the sparse
 * backward (0 without a block). */
int lapgpu_last_dense_timings(lapgpu_ctx* ctx, double* tail_fwd_ms, double* dense_ms,
                              double* tail_bwd_ms);

/* Average device time per call of each kernel class over the last solve
 * (ms), measured with HIP events when LAPGPU_PROFILE is set in flags. */
#define LAPGPU_PROFILE 4
int lapgpu_last_timings(lapgpu_ctx* ctx, double* spmv_ms, double* ldl_ms,
                        double* vec_ms, double* total_ms, int64_t* its);

/* Size of the level groups of the uploaded factor (the levels before the
 * tail): number of groups (launches per sweep), rows, and composite-row
 * entries of the forward and backward sweeps.  Measurement only
 * (algorithmic bytes of the group kernels). */
int lapgpu_group_info(lapgpu_ctx* ctx, int64_t* ngroups, int64_t* rows, int64_t* fentries,
                      int64_t* bentries);

/* Size of the single-workgroup tail of the uploaded factor: tail vertices,
 * program steps and slots of both sweeps, and forward in-edges from outside
 * the tail.  Measurement only (algorithmic bytes of the tail kernel). */
int lapgpu_tail_info(lapgpu_ctx* ctx, int64_t* nt, int64_t* nsteps, int64_t* nslots,
                     int64_t* nedges_in);

/* Autotuner (no reference counterpart; the reference has no device
 * schedule).  Times, on this graph, every level group's kernel with each
 * candidate lanes-per-row G in {4, 8, 16, 32} (rows longer than 8 G entries
 * get a wave each) and, when k > 1, each candidate slices-per-row S of the
 * k-column (rounded up to 8..64) multi-RHS kernels, and keeps the fastest
 * per group and sweep direction; with LAPGPU_TUNE_GROUPS it first tries
 * five level-group budgets (the context's depth and growth with 65536,
 * 196608 and 524288 extra entries, then 16 / 4 / 65536 and 8 / 2 / 16384,
 * whose many narrow groups the chained runs take; a regrouping each,
 * seconds on large graphs) and keeps the one with the fastest LDLinv apply.  Later solves use the
 * tuned schedule.  The lane split changes the order in which a row's terms
 * are summed, so tuned results may differ from untuned ones in the last
 * bits (and, through timing noise, between tuning runs).  times_ms (4
 * doubles, may be NULL): the device time of one single-RHS LDLinv apply
 * before and after, and of one k-column apply before and after (0 when
 * k <= 1). */
#define LAPGPU_TUNE_GROUPS 1
int lapgpu_autotune(lapgpu_ctx* ctx, int64_t k, int flags, double* times_ms);

/* The LDLinv apply of the last LAPGPU_PROFILE solve (single- or multi-RHS)
 * split by kernel (ms per iteration): forward level groups, the tail's
 * inputs (k_tail_ext / k_mtail_ext), the tail (k_tail / k_mtail) and
 * backward level groups.  Measurement only; no reference counterpart. */
int lapgpu_last_sweep_timings(lapgpu_ctx* ctx, double* fgroups_ms, double* tail_ext_ms,
                              double* tail_ms, double* bgroups_ms);

const char* lapgpu_last_error(lapgpu_ctx* ctx);

/* Release a context.  Device memory is stream-ordered (hipFreeAsync on the
 * context's stream, from the device's default pool, which keeps freed memory
 * for reuse) and pinned buffers return to a process-wide free list, so
 * neither creating nor destroying a context waits for other contexts' work
 * on the device. */
void lapgpu_destroy(lapgpu_ctx* ctx);

/* ---- batched solves: many graphs and many right-hand sides ------------- */
/* No counterpart in the reference, which solves one b per call; this is the
 * batched form of the same closure (src/approxChol.jl:1624) for callers that
 * hold many graphs or many right-hand sides (blockSolver's components,
 * src/solverInterface.jl:152-188; serving).  Each request runs the complete
 * pcg of src/pcg.jl:256-445 with LDLsolver in one workgroup, all requests in
 * one launch.  A batch may be used from several threads: every entry point
 * holds the batch's lock, except that lapgpu_batch_add_graph plans the graph
 * on the calling thread before it takes the lock (to upload and register
 * it), so graphs can be added while other threads submit and collect.
 * lapgpu_batch_last_error reports the calling thread's last failure only
 * when no other thread fails meanwhile. */
typedef struct lapgpu_batch lapgpu_batch;

int lapgpu_batch_create(lapgpu_batch** out, int device);

/* Upload one graph (la as in lapgpu_upload_csc, LDLinv as in
 * lapgpu_upload_ldli); *graph_id receives its handle. */
int lapgpu_batch_add_graph(lapgpu_batch* B, int64_t n, const int64_t* colptr,
                           const int64_t* rowval, const double* nzval,
                           int64_t nnzF, const int64_t* col,
                           const int64_t* fcolptr, const int64_t* frowval,
                           const double* fval, const double* d, int index_base,
                           int32_t* graph_id);

/* Free a graph's device memory.  Returns at once: the frees are ordered
 * after the uncollected launches that read the graph, and no other launch
 * is waited for.  A later lapgpu_batch_add_graph may reuse the id. */
int lapgpu_batch_remove_graph(lapgpu_batch* B, int32_t graph_id);

/* n and the device bytes held by a graph (for admission by HBM footprint). */
int lapgpu_batch_graph_info(lapgpu_batch* B, int32_t graph_id, int64_t* n,
                            int64_t* device_bytes);

/* Solve nreq requests; request i solves graph graph_ids[i] with the
 * right-hand side at b_concat[off_i .. off_i + n_i) (host buffers,
 * concatenated in request order), writing x_concat likewise.  Stopping
 * rules, tol/maxits/maxtime/stag_test/flags and the per-request outputs mean
 * what they mean in lapgpu_solve. */
int lapgpu_batch_solve(lapgpu_batch* B, int64_t nreq, const int32_t* graph_ids,
                       const double* b_concat, double* x_concat, double tol,
                       int64_t maxits, double maxtime, int stag_test, int flags,
                       int64_t* its_out, double* relres_out, int32_t* stop_out);

/* As lapgpu_batch_solve with per-request device pointers (resident b, x). */
int lapgpu_batch_solve_device(lapgpu_batch* B, int64_t nreq,
                              const int32_t* graph_ids,
                              const double* const* d_b, double* const* d_x,
                              double tol, int64_t maxits, double maxtime,
                              int stag_test, int flags, int64_t* its_out,
                              double* relres_out, int32_t* stop_out);

/* Asynchronous form: enqueue a launch of nreq requests (device pointers) on
 * one of the batch's launch slots and return at once; *slot_out names it.
 * Launches on different slots run concurrently on the device.  Returns
 * LAPGPU_BUSY when every slot holds an uncollected launch. */
#define LAPGPU_BUSY 1
int lapgpu_batch_submit(lapgpu_batch* B, int64_t nreq, const int32_t* graph_ids,
                        const double* const* d_b, double* const* d_x, double tol,
                        int64_t maxits, double maxtime, int stag_test, int flags,
                        int32_t* slot_out);
/* As lapgpu_batch_submit with the right-hand sides concatenated in a host
 * buffer (copied before the call returns); the solutions come back through
 * lapgpu_batch_collect's x_concat. */
int lapgpu_batch_submit_host(lapgpu_batch* B, int64_t nreq, const int32_t* graph_ids,
                             const double* b_concat, double tol, int64_t maxits,
                             double maxtime, int stag_test, int flags, int32_t* slot_out);
/* *done = 1 when the launch on `slot` has finished, else 0. */
int lapgpu_batch_query(lapgpu_batch* B, int32_t slot, int* done);
/* Wait for the launch on `slot`, write its per-request results (in the
 * order of the submit call) and free the slot.  x_concat receives the
 * solutions of a lapgpu_batch_submit_host launch (NULL otherwise). */
int lapgpu_batch_collect(lapgpu_batch* B, int32_t slot, double* x_concat, int64_t* its_out,
                         double* relres_out, int32_t* stop_out);

/* Cancel request `index` (its position in the submit call) of the launch
 * on `slot` (no reference counterpart; for serving): its workgroup stops at
 * its next iteration boundary with stop reason LAPGPU_CANCELLED and its
 * best iterate so far; every other request of the launch runs on unchanged
 * (each has its own workgroup and nothing is shared).  The flag is read by
 * the device from host memory while the launch runs; a request that has
 * already finished keeps its result.  The slot must hold the launch the
 * caller means (between its submit and its collect). */
int lapgpu_batch_cancel(lapgpu_batch* B, int32_t slot, int64_t index);

/* Allocation-free form of lapgpu_batch_solve_device (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as code
 * under HIP graph capture, tight serving loops and harnesses that forbid
 * hipMalloc.  It calls no hipMalloc, hipFree or pinned allocation: the
 * requests' work vectors (8 n doubles each, plus up to 1024 partial
 * slots), request table and graph
 * descriptors live in a static device array of 64 MiB per device (about
 * 1,000,000 vertices over all requests of one call), and at most 1024
 * requests go in one call; larger calls fail with a message.  Same kernel
 * and results as lapgpu_batch_solve_device; synchronous.  Calls share the
 * device's array and are serialised by a process-wide lock; the array keeps
 * the last call's contents. */
int lapgpu_batch_solve_device_noalloc(lapgpu_batch* B, int64_t nreq, const int32_t* graph_ids,
                                      const double* const* d_b, double* const* d_x, double tol,
                                      int64_t maxits, double maxtime, int stag_test, int flags,
                                      int64_t* its_out, double* relres_out, int32_t* stop_out);

const char* lapgpu_batch_last_error(lapgpu_batch* B);
void lapgpu_batch_destroy(lapgpu_batch* B);

#ifdef __cplusplus
}
#endif

#endif /* LAPGPU_H */
