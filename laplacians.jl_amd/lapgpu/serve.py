"""Serving loop: continuous batching of Laplacian solves on one GPU.

Graphs (la + LDLinv) stay resident in HBM.  Requests (graph, b, tol) arrive
over time; each is answered by the same solve closure as approxchol_lap's
(src/approxChol.jl:1624) on the device.  The loop

* admits a request as soon as it arrives: graphs of at most `wg_max_n`
  vertices (`new_wg_max_n` for a graph that comes with its request) are
  queued for the next batched launch (one workgroup per request,
  lapgpu_batch_submit_host, whose workgroups start largest graph first);
  larger graphs queue per graph, and a
  worker solves everything queued for that graph in one multi-RHS solve
  (lapgpu_solve_multi, up to the graph's `max_rhs` columns sharing every
  launch; lapgpu_solve when only one is queued) on the graph's own stream;
  all kinds overlap on the device.  The batch is taken when the worker
  starts, and an idle graph waits up to half its last solve time (at most
  `max_wait`) for more requests unless `max_rhs` are already queued;
* keeps up to four batched launches in flight (the batch's launch slots):
  requests that arrive while others run start in the next free slot instead
  of waiting for the running launch to finish;
* admits a new graph by its HBM footprint: when the budget is exceeded,
  idle non-resident graphs are evicted (least recently used first), and a
  request waits while its graph cannot be admitted;
* factors new graphs on host threads (liblapfactor, re-entrant) while the
  device keeps solving; the same threads plan and upload it (to the batched
  engine, or to a context of its own when it is larger; its HBM is reserved
  first), and one more thread closes dropped
  contexts, so the dispatch loop itself never waits on an upload or a free.
  A graph that comes with its request is planned with a small level-group
  budget (`new_graph_budget`): fewer composite rows to build on the host, a
  few more launches on the device for its one solve; and without a dense
  tail block (lapgpu_set_dense_tail), whose matrix costs more to build than
  one solve saves.  Registered graphs' contexts get a small budget too
  (`resident_budget`), for their multi-RHS solves.

Every result carries its latency (arrival to completion); `summary` gives
solves/s and p50/p99 latency.  `cancel(rid)` withdraws a request from any
thread: answered at once while it is queued, stopped at its next iteration
boundary once it is on the device, without changing the requests that share
its launch or multi-RHS solve.

Every graph context owns a stream, and HIP maps streams round-robin onto its
hardware queues (4 unless GPU_MAX_HW_QUEUES says otherwise): a small solve on
a stream that shares a queue with a long multi-RHS solve waits for it.  A
process that serves many graphs should set GPU_MAX_HW_QUEUES=16 before its
first HIP call (more than 16 queues slowed plain solves on MI355X).
"""

from __future__ import annotations

import threading
import time
from collections import OrderedDict, deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from . import _native
from .cache import GraphCache
from .factor import LDLinv, build_ldli
from .graphs import lap


@dataclass
class Request:
    rid: int
    arrival: float  # seconds after the start of the run
    b: np.ndarray
    graph: object = None  # key of a registered graph
    new_graph: sp.spmatrix | None = None  # adjacency that needs factoring first
    tol: float = 1e-8
    maxits: int = 1000


@dataclass
class Result:
    rid: int
    latency: float
    its: int
    nr: float
    stop: str
    path: str
    x: np.ndarray | None = None


@dataclass
class _Graph:
    key: object
    n: int
    kind: str  # "wg" or "plain"
    bytes: int
    resident: bool
    handle: object = None  # batch graph id or Context
    busy: int = 0
    # replicas: contexts holding the same graph (handle is the first), so
    # that as many multi-RHS solves of it run at once; idle: those free;
    # source: (la, LDLinv, cache key) of a registered graph, for add_replicas
    replicas: list = field(default_factory=list)
    idle: list = field(default_factory=list)
    source: tuple | None = None
    pending: deque = field(default_factory=deque)
    max_rhs: int = 1  # columns per multi-RHS solve on this graph
    last_solve_s: float = 0.0  # duration of the graph's last solve
    ready_at: float = 0.0  # an idle graph with few requests waits until then
    solves: int = 0  # plain / multi-RHS solves run on the graph's context


DENSE_MAX = 18432  # the largest dense tail block a context may hold (its LDS tail)


def dense_bytes_estimate(n: int) -> int:
    """The largest dense tail block a context of n vertices can hold (8 K^2
    bytes, K <= min(n, 18432)); the planner usually chooses less, and the
    reservation is corrected once the context reports its block."""
    k = min(n, DENSE_MAX)
    return 8 * k * ((k + 255) // 256 * 256)


def graph_bytes_estimate(n: int, nnz_l: int, nnzf: int, fp32: bool = False) -> int:
    """Device bytes of a graph for admission: la (12 B/entry), the level-
    ordered factor in both sweep directions (about 16 B/entry each) and 8
    workspace vectors (a context adds its dense tail block,
    dense_bytes_estimate); a context in the FP32-inner mode adds FP32 copies
    of la's values and the sweep weights and 20 B per vertex of FP32/FP64
    work vectors."""
    b = 12 * nnz_l + 32 * nnzf + 8 * 8 * n + 16 * n
    if fp32:
        b += 4 * nnz_l + 8 * nnzf + 20 * n
    return b


def dense_correction(ctx, n: int) -> int:
    """Actual minus reserved bytes of a context's dense tail block."""
    return ctx.dense_info()["bytes"] - dense_bytes_estimate(n)


def multi_bytes_estimate(n: int, k: int, fp32: bool = False) -> int:
    """Workspace of a k-column multi-RHS solve: 9 interleaved n-by-K vectors
    (K = k rounded up to a power of two, at least 8), plus 20 B per vertex and
    column in the FP32-inner mode (y, z and p in FP32, r0 in FP64)."""
    K = 8
    while K < k:
        K *= 2
    return (9 * 8 + (20 if fp32 else 0)) * n * K


class Server:
    def __init__(self, device: int = 0, hbm_budget: int | None = None, wg_max_n: int = 50_000,
                 max_launch: int = 512, plain_workers: int = 4, factor_workers: int = 4,
                 keep_x: bool = False, max_rhs: int = 64, batch_window: float = 0.5,
                 max_wait: float = 1.0, new_graph_budget: tuple | None = (8, 2.0, 16384),
                 resident_budget: tuple | None = (16, 4.0, 65536),
                 eager_new: bool = True, quick_workers: int = 4, quick_max_n: int = 300_000,
                 blocking_waits: bool = True, keep_ids: set | None = None,
                 new_wg_max_n: int | None = 200_000, cache_dir: str | None = None,
                 context_precision: str = "fp64", exact_batches: bool = False,
                 replica_batch: int = 16, idle_sleep: float = 5e-4):
        self.device = device
        # the dispatch loop's pause after a pass that changed nothing: it
        # holds the interpreter lock while it scans, which the solve workers
        # and the factoring threads need between their native calls
        # (measured on the serving stream: 0.5 ms against 50 us, host CPU
        # 389-405 against 444 core-seconds, throughput the same within 2 %)
        self.idle_sleep = idle_sleep
        # factors and context plans of graphs seen before (lapgpu.cache):
        # a graph that comes back is admitted without factoring or planning
        self.cache = GraphCache(cache_dir) if cache_dir else None
        # contexts (graphs above the batched engine's bound) solve in FP64 or
        # in the FP32-inner mode; either way a request converges only on its
        # FP64 true residual, so every answer meets its tol
        if context_precision not in ("fp64", "fp32"):
            raise ValueError(f"context_precision must be 'fp64' or 'fp32', not {context_precision!r}")
        self.context_fp32 = context_precision == "fp32"
        # the solutions kept in the results: all (keep_x) or those of keep_ids
        self.keep_ids = keep_ids
        self.batch = _native.Batch(device)
        self.hbm_budget = hbm_budget if hbm_budget is not None else 200 * 2**30
        self.wg_max_n = wg_max_n
        # a graph that comes with its request is solved once: on the batched
        # engine up to new_wg_max_n vertices (one CU for the solve, and no
        # context, stream or level groups to set up on the host)
        self.new_wg_max_n = max(wg_max_n, new_wg_max_n or 0)
        self.max_launch = max_launch
        self.keep_x = keep_x
        self.max_rhs = max_rhs
        # an idle graph whose first queued request has waited less than
        # batch_window x its last solve time (at most max_wait seconds) and
        # that has fewer than max_rhs requests queued collects more first
        self.batch_window = batch_window
        self.max_wait = max_wait
        # level-group budget (lapgpu_set_group_budget) of graphs that come
        # with a request and are dropped after it: planning composite rows
        # costs host time in proportion to their entries, and a graph solved
        # once recovers little of it (None: the library's default budget)
        self.new_graph_budget = new_graph_budget
        # level-group budget of registered graphs' contexts, which serve
        # multi-RHS solves: every composite-row entry is read once per
        # column there, so fewer of them pay (a 3.4M-vertex path + matchings
        # graph's 64-column solve alone: 2.85 s with the library's default
        # budget, 1.73 s with 8 / 2 / 16,384), but beside many other solves
        # each dependent launch also waits its turn on the device, so fewer
        # launches pay too: 16 / 4 / 65,536 served a mixed stream fastest
        # (236 solves/s against 221 with 8 / 2 / 16,384 and 159 with the
        # library's default)
        self.resident_budget = resident_budget
        self.eager_new = eager_new
        # multi-RHS batches of exactly 8, 16, 32 or 64 requests (K, the
        # columns every launch of the solve processes, padding included),
        # the rest waiting for the graph's next solve: on a saturated device
        # a padding column costs as much as a request
        self.exact_batches = exact_batches
        # a graph with replicas (add_replicas) starts a second solve while
        # one runs once replica_batch requests are queued
        self.replica_batch = replica_batch
        self._take_lock = threading.Lock()
        # solve workers sleep while the device works (lapgpu_set_wait_mode)
        # instead of spinning, leaving the cores to factoring and planning
        self.blocking_waits = blocking_waits
        # (n, columns, seconds, first on its context, CPU seconds of the
        # worker thread) of each plain / multi-RHS solve
        self.solve_log: list = []
        # seconds the dispatch loop spent admitting and dropping graphs
        self.stats = {"admit_s": 0.0, "drop_s": 0.0, "admits": 0, "drops": 0, "loop_cpu_s": 0.0}
        # worker seconds per new graph: factoring; context upload (total,
        # create, Laplacian, factor)
        self.factor_log: list = []
        self.factor_cpu_log: list = []
        self.upload_cpu_log: list = []
        self.upload_log: list = []
        # (seconds, CPU seconds) of each batched-engine graph upload
        self.wg_upload_log: list = []
        self.graphs: "OrderedDict[object, _Graph]" = OrderedDict()
        self.hbm_used = 0
        self._plain_pool = ThreadPoolExecutor(max_workers=plain_workers)
        # solves of graphs with at most quick_max_n vertices (tens of ms) get
        # workers of their own, so they never queue behind the multi-second
        # multi-RHS solves of large graphs
        self.quick_max_n = quick_max_n
        self._quick_pool = ThreadPoolExecutor(max_workers=quick_workers) if quick_workers > 0 else None
        self._factor_pool = ThreadPoolExecutor(max_workers=factor_workers)
        self._close_pool = ThreadPoolExecutor(max_workers=1)
        self._anon = 0
        # cancellation (cancel(), from any thread): every rid ever cancelled,
        # and those whose cancel the dispatch loop has not carried out yet
        self._cancel_lock = threading.Lock()
        self._cancelled: set = set()
        self._cancel_todo: set = set()
        # rid -> (graph, column) of the requests in a running context solve,
        # rid -> (launch handle, index) of those in a batched launch
        self._running: dict = {}
        self._in_launch: dict = {}

    def cancel(self, rid) -> None:
        """Withdraw request `rid` (thread-safe; for a running run()).  A
        request still queued, factoring or waiting for admission is answered
        at once with stop "cancelled" and no solution; one already on the
        device stops at its next iteration boundary (lapgpu_batch_cancel /
        lapgpu_cancel) with stop "cancelled" and its best iterate, while the
        requests sharing its launch or multi-RHS solve run on unchanged.  A
        request that has finished keeps its result."""
        with self._cancel_lock:
            self._cancelled.add(rid)
            self._cancel_todo.add(rid)

    def on_device(self, rid) -> bool:
        """True while request `rid` is in a running batched launch or
        context solve (thread-safe)."""
        return rid in self._in_launch or rid in self._running

    def _is_cancelled(self, rid) -> bool:
        with self._cancel_lock:
            return rid in self._cancelled

    # ------------------------------------------------------------ graphs ---
    def register(self, key, a: sp.spmatrix, ldli: LDLinv | None = None, seed: int = 1,
                 resident: bool = True) -> bool:
        """Upload a graph (factoring it on the host if no LDLinv is given).
        Returns False when it does not fit in the HBM budget."""
        a = sp.csc_matrix(a)
        gkey = None
        if ldli is None and self.cache is not None:
            gkey, la, ldli, _ = self.cache.factor(a, seed=seed)
        else:
            la = lap(a)
            if ldli is None:
                ldli = build_ldli(a, seed=seed)
        return self._admit(key, la, ldli, resident, gkey)

    def _reserve(self, la, ldli, wg: bool):
        """HBM footprint and multi-RHS width of a graph on the batched engine
        (wg) or its own context, evicting idle non-resident graphs when
        needed; None when it does not fit."""
        n = la.shape[0]
        need = graph_bytes_estimate(n, la.nnz, ldli.nnz, fp32=self.context_fp32 and not wg)
        if not wg:
            need += dense_bytes_estimate(n)  # contexts only (not the batched engine)
        if self.hbm_used + need > self.hbm_budget:
            self._evict(need)
            if self.hbm_used + need > self.hbm_budget:
                return None
        k = 1
        if not wg:
            # the widest multi-RHS workspace that still fits the budget
            k = self.max_rhs
            fp32 = self.context_fp32
            while k > 1 and self.hbm_used + need + multi_bytes_estimate(n, k, fp32) > self.hbm_budget:
                k //= 2
            if k > 1:
                need += multi_bytes_estimate(n, k, fp32)
        return need, k

    def _upload_context(self, la, ldli, budget=None, dense_max=None, gkey=None):
        c0 = time.thread_time()
        t0 = time.perf_counter()
        ctx = _native.Context(self.device)
        if self.blocking_waits:
            ctx.set_wait_mode(True)
        if budget is not None:
            ctx.set_group_budget(*budget)
        if dense_max is not None:
            ctx.set_dense_tail(dense_max)
        t1 = time.perf_counter()
        if self.cache is not None and gkey is not None:
            # the plan depends on the settings it was built with
            settings = f"b{'-'.join(map(str, budget)) if budget else 'default'}_d{dense_max}"
            self.cache.upload(ctx, gkey, settings, la, ldli)
            t2 = time.perf_counter()
        else:
            ctx.upload_lap(la)
            t2 = time.perf_counter()
            ctx.upload_ldli(ldli)
        if self.context_fp32:
            ctx.set_precision(_native.FP32_INNER, 20)
        t3 = time.perf_counter()
        self.upload_log.append((t3 - t0, t1 - t0, t2 - t1, t3 - t2))
        self.upload_cpu_log.append(time.thread_time() - c0)
        return ctx

    def _upload_wg(self, la, ldli):
        """Plan and upload a graph to the batched engine (thread-safe:
        add_graph plans outside the batch's lock); returns its id."""
        c0 = time.thread_time()
        t0 = time.perf_counter()
        gid = self.batch.add_graph(la, ldli)
        self.wg_upload_log.append((time.perf_counter() - t0, time.thread_time() - c0))
        return gid

    def _wg_graph(self, key, n, gid, resident) -> _Graph:
        g = _Graph(key, n, "wg", self.batch.graph_info(gid)["device_bytes"] + 64 * n, resident, gid)
        self.graphs[key] = g
        self.hbm_used += g.bytes
        return g

    def _admit(self, key, la, ldli, resident, gkey=None) -> bool:
        n = la.shape[0]
        wg = n <= self.wg_max_n
        res = self._reserve(la, ldli, wg)
        if res is None:
            return False
        need, k = res
        if wg:
            self._wg_graph(key, n, self._upload_wg(la, ldli), resident)
        else:
            ctx = self._upload_context(la, ldli, budget=self.resident_budget, gkey=gkey)
            g = _Graph(key, n, "plain", need + dense_correction(ctx, n), resident, ctx, max_rhs=k)
            g.replicas = [ctx]
            g.idle = [ctx]
            if resident:
                g.source = (la, ldli, gkey)
            self.graphs[key] = g
            self.hbm_used += g.bytes
        return True

    def add_replicas(self, min_nnzf: int, copies: int = 2) -> list:
        """Give registered graphs on contexts whose LDLinv has at least
        min_nnzf entries up to `copies` contexts each, largest factors first,
        while the HBM budget allows (call between runs, after registering).
        Their multi-RHS solves are the longest: one context (one solve at a
        time) falls behind once requests arrive faster than a solve of the
        batch they form completes, and a second context lets two such solves
        run at once.  Returns the keys that got a replica."""
        got = []
        cands = [g for g in self.graphs.values()
                 if g.kind == "plain" and g.resident and g.source is not None and g.source[1].nnz >= min_nnzf]
        for g in sorted(cands, key=lambda g: -g.source[1].nnz):
            per = g.bytes // len(g.replicas)
            while len(g.replicas) < copies and self.hbm_used + per <= self.hbm_budget:
                la, ldli, gkey = g.source
                ctx = self._upload_context(la, ldli, budget=self.resident_budget, gkey=gkey)
                g.replicas.append(ctx)
                g.idle.append(ctx)
                g.bytes += per
                self.hbm_used += per
                got.append(g.key)
        return got

    def _evict(self, need: int) -> None:
        for key in list(self.graphs):
            if self.hbm_used + need <= self.hbm_budget:
                return
            g = self.graphs[key]
            if g.resident or g.busy or g.pending:
                continue
            self._drop(g)

    def _drop(self, g: _Graph) -> None:
        if g.kind == "wg":
            self.batch.remove_graph(g.handle)
        else:
            # closing waits for the device's frees; the contexts are idle
            for ctx in g.replicas or [g.handle]:
                self._close_pool.submit(ctx.close)
        self.hbm_used -= g.bytes
        del self.graphs[g.key]

    # --------------------------------------------------------------- run ---
    def run(self, requests: list[Request], realtime: bool = True, progress=None,
            progress_every: float = 20.0) -> list[Result]:
        """Serve `requests` (sorted by arrival).  With realtime=False every
        request is taken to have arrived at time 0 (throughput mode).
        progress(done, total, seconds), if given, is called by the dispatch
        loop every progress_every seconds."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        results: dict[int, Result] = {}
        arrival_abs = {}
        small: deque = deque()  # (graph, request) ready for a batched launch
        launches = []  # (handle, [(graph, request)])
        plain_futs = {}
        factor_futs = {}
        upload_futs = {}  # context uploads of admitted large graphs
        waiting_admission: deque = deque()
        waiting: set = set()  # idle graphs collecting a batch
        in_launch = self._in_launch  # rid -> (launch handle, index in the launch)
        in_launch.clear()
        placed = {}  # rid -> graph it was routed to
        t0 = time.perf_counter()
        c0 = time.thread_time()
        nxt = 0
        total = len(reqs)
        next_progress = progress_every

        def finish(r: Request, info: dict, x, path: str, column: bool = False):
            # column: x is a column view of a multi-RHS solution block, copied
            # only when the result is kept (a copy of every column of a
            # multi-million-vertex block on this loop delays every dispatch)
            lat = time.perf_counter() - arrival_abs[r.rid]
            keep = self.keep_x or (self.keep_ids is not None and r.rid in self.keep_ids)
            if keep and column:
                x = x.copy()
            results[r.rid] = Result(r.rid, lat, info["its"], info["nr"], info["stop"], path, x if keep else None)

        def finish_cancelled(r: Request, path: str):
            finish(r, {"its": 0, "nr": 1.0, "stop": "cancelled"}, None, path)

        def route(g: _Graph, r: Request):
            if self._is_cancelled(r.rid):
                finish_cancelled(r, "queued")
                return
            placed[r.rid] = g
            if g.kind == "wg":
                g.busy += 1
                small.append((g, r))
            else:
                if not g.pending and not g.busy:
                    g.ready_at = time.perf_counter() + min(self.batch_window * g.last_solve_s,
                                                           self.max_wait)
                g.pending.append(r)
                pump_plain(g)

        def solve_queued(g: _Graph, ctx):
            # runs in a worker on one of the graph's contexts: takes
            # everything queued for the graph NOW (not when the task was
            # submitted, so requests that arrive while the task waits for a
            # worker join it), with the same tol and maxits as the oldest, up
            # to max_rhs columns
            with self._take_lock:  # replicas' workers take from one queue
                take = [g.pending.popleft()] if g.pending else []
                while (g.pending and len(take) < g.max_rhs and g.pending[0].tol == take[0].tol
                       and g.pending[0].maxits == take[0].maxits):
                    take.append(g.pending.popleft())
                if self.exact_batches and len(take) > 8:
                    # a multi-RHS solve runs K = a power of two columns: take
                    # that many and leave the rest, first in line, for the
                    # next solve
                    m = 1 << (len(take).bit_length() - 1)
                    for r in reversed(take[m:]):
                        g.pending.appendleft(r)
                    take = take[:m]
            # requests cancelled while queued are answered without a solve
            gone = [r for r in take if self._is_cancelled(r.rid)]
            take = [r for r in take if not self._is_cancelled(r.rid)]
            if not take:
                return take, gone, None, None
            for j, r in enumerate(take):
                self._running[r.rid] = (ctx, j)
            r0 = take[0]
            # a graph that came with its request is solved without capturing
            # its iteration as a HIP graph (the capture would be used once)
            flags = _native.EAGER if (self.eager_new and not g.resident) else 0
            c_start = time.thread_time()
            t_start = time.perf_counter()
            if len(take) == 1:
                x, info = ctx.solve(r0.b, r0.tol, r0.maxits, float("inf"), 5, flags)
            else:
                # column-major, as the library reads it: filled column by
                # column (stacking row-major and converting moved every
                # element twice with a stride: seconds for a 64-column block
                # of a 3.4M-vertex graph)
                B = np.empty((len(r0.b), len(take)), order="F")
                for j, r in enumerate(take):
                    B[:, j] = r.b
                x, info = ctx.solve_multi(B, r0.tol, r0.maxits, float("inf"), 5, flags)
            dt = time.perf_counter() - t_start
            for r in take:
                self._running.pop(r.rid, None)
            self.solve_log.append((g.n, len(take), dt, g.solves == 0, time.thread_time() - c_start))
            g.solves += 1
            g.last_solve_s = dt
            return take, gone, x, info

        def pump_plain(g: _Graph):
            # one solve at a time per context (a context owns one stream); a
            # graph with replicas starts another solve while one runs once
            # replica_batch requests are queued
            if not g.pending:
                waiting.discard(g.key)
            idle = g.idle if g.replicas else ([g.handle] if not g.busy else [])
            if not idle or not g.pending:
                return
            if g.busy:
                if len(g.pending) < min(g.max_rhs, self.replica_batch):
                    return
            elif len(g.pending) < g.max_rhs and time.perf_counter() < g.ready_at:
                waiting.add(g.key)
                return
            waiting.discard(g.key)
            ctx = idle.pop(0) if g.replicas else g.handle
            g.busy += 1
            self.stats["max_busy_contexts"] = max(self.stats.get("max_busy_contexts", 0), g.busy)
            quick = self._quick_pool is not None and g.n <= self.quick_max_n
            pool = self._quick_pool if quick else self._plain_pool
            plain_futs[pool.submit(solve_queued, g, ctx)] = (g, ctx)

        while len(results) < total:
            now = time.perf_counter() - t0
            progressed = False
            # arrivals
            while nxt < total and (not realtime or reqs[nxt].arrival <= now):
                r = reqs[nxt]
                nxt += 1
                arrival_abs[r.rid] = t0 + (r.arrival if realtime else 0.0)
                if self._is_cancelled(r.rid):
                    finish_cancelled(r, "queued")
                elif r.new_graph is not None:
                    fut = self._factor_pool.submit(self._factor, r.new_graph)
                    factor_futs[fut] = r
                else:
                    route(self.graphs[r.graph], r)
                progressed = True
            # factored graphs: admit by HBM footprint, then route
            for fut in [f for f in factor_futs if f.done()]:
                r = factor_futs.pop(fut)
                waiting_admission.append((r, fut.result()))
                progressed = True
            for _ in range(len(waiting_admission)):
                r, (la, ldli, gkey) = waiting_admission.popleft()
                if self._is_cancelled(r.rid):
                    finish_cancelled(r, "queued")
                    progressed = True
                    continue
                ta = time.perf_counter()
                wg = la.shape[0] <= self.new_wg_max_n
                res = self._reserve(la, ldli, wg)
                if res is None:
                    waiting_admission.append((r, (la, ldli, gkey)))
                    continue
                self._anon += 1
                key = ("new", self._anon)
                need, k = res
                # reserved now (a batched-engine graph by the estimate until
                # it reports its own bytes), uploaded on a worker thread
                self.hbm_used += need
                if wg:
                    fut = self._factor_pool.submit(self._upload_wg, la, ldli)
                else:
                    fut = self._factor_pool.submit(self._upload_context, la, ldli, self.new_graph_budget, 0,
                                                   gkey)
                upload_futs[fut] = (key, r, la.shape[0], need, k, wg)
                self.stats["admit_s"] += time.perf_counter() - ta
                self.stats["admits"] += 1
                progressed = True
            for fut in [f for f in upload_futs if f.done()]:
                key, r, n, need, k, wg = upload_futs.pop(fut)
                if wg:
                    self.hbm_used -= need
                    g = self._wg_graph(key, n, fut.result(), False)
                else:
                    ctx = fut.result()
                    corr = dense_correction(ctx, n)
                    g = _Graph(key, n, "plain", need + corr, False, ctx, max_rhs=k)
                    self.hbm_used += corr
                    self.graphs[key] = g
                route(g, r)
                progressed = True
            # idle graphs whose batching window has ended
            for key in list(waiting):
                g = self.graphs.get(key)
                if g is None:
                    waiting.discard(key)
                    continue
                if len(g.pending) >= g.max_rhs or time.perf_counter() >= g.ready_at:
                    pump_plain(g)
                    progressed = True
            # finished batched launches
            for item in [it for it in launches if self.batch.query(it[0])]:
                launches.remove(item)
                xs, infos = self.batch.collect(item[0])
                for (g, r), x, info in zip(item[1], xs, infos):
                    g.busy -= 1
                    in_launch.pop(r.rid, None)
                    finish(r, info, x, "batched")
                progressed = True
            # finished plain solves
            for fut in [f for f in plain_futs if f.done()]:
                g, ctx = plain_futs.pop(fut)
                take, gone, x, info = fut.result()
                g.busy -= 1
                if g.replicas:
                    g.idle.append(ctx)
                g.ready_at = 0.0  # requests queued meanwhile go at once
                for r in gone:
                    finish_cancelled(r, "queued")
                if not take:
                    pass
                elif len(take) == 1:
                    finish(take[0], info, x, "plain")
                else:
                    for j, r in enumerate(take):
                        finish(r, {"its": info["its"][j], "nr": info["nr"][j], "stop": info["stop"][j]},
                               x[:, j], "multi", column=True)
                pump_plain(g)
                progressed = True
            # cancellations asked for since the last pass
            with self._cancel_lock:
                todo = list(self._cancel_todo)
            for rid in todo:
                done_now = rid in results
                if not done_now and rid in in_launch:
                    h, i = in_launch[rid]
                    self.batch.cancel(h, i)
                    done_now = True
                elif not done_now and rid in self._running:
                    ctx, j = self._running.get(rid, (None, 0))
                    # False: the solve has not started or has just ended
                    done_now = ctx is not None and ctx.cancel(j)
                elif not done_now and rid in placed:
                    g = placed[rid]
                    if g.kind == "wg":
                        for item in list(small):
                            if item[1].rid == rid:
                                small.remove(item)
                                g.busy -= 1
                                finish_cancelled(item[1], "queued")
                                done_now = True
                    elif not g.busy:
                        # no worker is taking from an idle graph's queue
                        for r in list(g.pending):
                            if r.rid == rid:
                                g.pending.remove(r)
                                finish_cancelled(r, "queued")
                                done_now = True
                if done_now:
                    with self._cancel_lock:
                        self._cancel_todo.discard(rid)
                    progressed = True
            # start a batched launch with everything queued, if a slot is free
            while small:
                # one launch runs every request with the same tol and maxits:
                # take the queued requests that share the oldest one's, in
                # arrival order, and leave the others queued
                key = (small[0][1].tol, small[0][1].maxits)
                take, rest = [], deque()
                while small and len(take) < self.max_launch:
                    item = small.popleft()
                    (take if (item[1].tol, item[1].maxits) == key else rest).append(item)
                rest.extend(small)
                small.clear()
                small.extend(rest)
                h = self.batch.submit_host([g.handle for g, _ in take], [r.b for _, r in take],
                                           tol=key[0], maxits=key[1])
                if h is None:
                    for item in reversed(take):
                        small.appendleft(item)
                    break
                launches.append((h, take))
                for i, (_, r) in enumerate(take):
                    in_launch[r.rid] = (h, i)
                progressed = True
            # drop finished one-off graphs
            for key in [k for k, g in self.graphs.items() if not g.resident and not g.busy and not g.pending]:
                td = time.perf_counter()
                self._drop(self.graphs[key])
                self.stats["drop_s"] += time.perf_counter() - td
                self.stats["drops"] += 1
            if progress is not None and time.perf_counter() - t0 >= next_progress:
                progress(len(results), total, time.perf_counter() - t0)
                next_progress += progress_every
            if not progressed:
                time.sleep(self.idle_sleep)
        self.stats["loop_cpu_s"] += time.thread_time() - c0
        return [results[r.rid] for r in requests]

    def _factor(self, a):
        """(la, LDLinv, cache key or None) of a graph that came with its
        request: from the cache when it has the graph, else factored."""
        c = time.thread_time()
        t = time.perf_counter()
        a = sp.csc_matrix(a)
        if self.cache is not None:
            gkey, la, f, _ = self.cache.factor(a, seed=1)
            out = la, f, gkey
        else:
            out = lap(a), build_ldli(a, seed=1), None
        self.factor_log.append(time.perf_counter() - t)
        self.factor_cpu_log.append(time.thread_time() - c)
        return out

    def close(self):
        self._plain_pool.shutdown(wait=True)
        if self._quick_pool is not None:
            self._quick_pool.shutdown(wait=True)
        self._factor_pool.shutdown(wait=True)
        self._close_pool.shutdown(wait=True)
        for g in list(self.graphs.values()):
            if g.kind == "plain":
                for ctx in g.replicas or [g.handle]:
                    ctx.close()
        self.graphs.clear()
        self.batch.close()


def summary(results: list[Result], wall: float) -> dict:
    """solves/s and latency percentiles over the results (cancelled
    requests included); max_nr over the requests that were not cancelled."""
    lat = np.array([r.latency for r in results])
    nrs = [r.nr for r in results if r.stop != "cancelled"]
    return {
        "solves": len(results),
        "wall_s": wall,
        "solves_per_s": len(results) / wall if wall > 0 else None,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "max_nr": float(max(nrs)) if nrs else None,
        "stops": sorted({r.stop for r in results}),
        "cancelled": sum(r.stop == "cancelled" for r in results),
    }
