"""The host planner (laplacians.jl_amd/csrc/lapgpu_plan.cpp) gives the same
plan, bit for bit, for any number of planner threads: every plan array
(level order, permuted la, both sweep layouts, level-group composite rows,
tail slot programs) is hashed by a small driver built from the planner's
source with g++ and compared between 1 and 4 threads."""

import os
import subprocess

import numpy as np
import pytest

import lapgpu
from lapgpu import graphs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "laplacians.jl_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("plan") / "plan_hash")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{ROOT}/include", f"-I{CSRC}", "-o", exe,
                    os.path.join(ROOT, "tests", "native", "plan_hash.cpp"), os.path.join(CSRC, "lapgpu_plan.cpp")],
                   check=True)
    return exe


def _dump(a, d):
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    os.makedirs(d, exist_ok=True)
    for k, v in (("cp", la.indptr), ("rv", la.indices), ("nz", la.data), ("col", f.col), ("fcp", f.colptr),
                 ("frv", f.rowval), ("fv", f.fval), ("d", f.d)):
        np.ascontiguousarray(v, dtype=np.float64 if k in ("nz", "fv", "d") else np.int64).tofile(os.path.join(d, k))


@pytest.mark.parametrize("name", ["grid2", "grid3", "path_matchings"])
def test_plan_same_for_any_thread_count(driver, tmp_path, name):
    rng = np.random.default_rng(11)
    a = {"grid2": lambda: graphs.weighted_grid2(300, 200, rng, "expU2"),
         "grid3": lambda: graphs.weighted_grid3(40, 40, 40, rng, "expU2"),
         "path_matchings": lambda: graphs.path_matchings(60000, 4, rng, "logU3")}[name]()
    d = str(tmp_path / name)
    _dump(a, d)
    outs = []
    for t in (1, 4):
        r = subprocess.run([driver, d], env={**os.environ, "LAPGPU_PLAN_THREADS": str(t)},
                           capture_output=True, text=True, check=True)
        outs.append(r.stdout)
    assert outs[0] == outs[1]
    # the grid cases need level groups before the tail to cover the composite rows
    last = outs[0].strip().splitlines()[-1].split()
    assert int(last[1]) > 1 and int(last[3]) > 0
    if name != "path_matchings":
        assert int(last[3]) > 1, "expected several level groups"


def test_entries_per_lane_follow_row_length(driver, tmp_path):
    """The tail program packs 2^elog entries per lane, chosen from its mean
    row length: 4 for long rows (a dense random graph, ~150 entries per
    row), 1 for a 2-D grid (~4)."""
    rng = np.random.default_rng(12)
    for name, a, want in (("dense", graphs.random_dense(1500, 120, rng), 2),
                          ("grid2", graphs.weighted_grid2(60, 50, rng, "expU2"), 0)):
        d = str(tmp_path / name)
        _dump(a, d)
        r = subprocess.run([driver, d], capture_output=True, text=True, check=True)
        elog = int(r.stdout.strip().splitlines()[-1].split()[7])
        assert elog == want, name


def test_root_scale_follows_its_d(driver, tmp_path):
    """LDLsolver divides y[root] by d[root] when it is nonzero
    (src/approxChol.jl:1351-1355).  The builders leave d[root] = 0 (scale 1);
    a caller's LDLinv with d[root] = 2.5 must give the root the scale
    1/2.5 = 0.4, and the root keeps scale 1 when d[root] = 0."""
    import dataclasses

    a = graphs.weighted_grid2(20, 15, np.random.default_rng(13), "expU2")
    f = lapgpu.build_ldli(a, seed=1)
    root = int(np.setdiff1d(np.arange(a.shape[0]), f.col)[0])
    assert f.d[root] == 0.0
    for droot, want in ((0.0, 1.0), (2.5, 0.4)):
        d = f.d.copy()
        d[root] = droot
        g = dataclasses.replace(f, d=d)
        out = str(tmp_path / f"root_{droot}")
        la = lapgpu.lap(a)
        os.makedirs(out, exist_ok=True)
        for k, v in (("cp", la.indptr), ("rv", la.indices), ("nz", la.data), ("col", g.col), ("fcp", g.colptr),
                     ("frv", g.rowval), ("fv", g.fval), ("d", g.d)):
            np.ascontiguousarray(v, dtype=np.float64 if k in ("nz", "fv", "d") else np.int64).tofile(
                os.path.join(out, k))
        r = subprocess.run([driver, out, "root_s"], capture_output=True, text=True, check=True)
        words = r.stdout.split()
        assert int(words[1]) == root
        assert float(words[3]) == want, (droot, r.stdout)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("prog") / "program_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{ROOT}/include", f"-I{CSRC}", "-o", exe,
                    os.path.join(ROOT, "tests", "native", "program_check.cpp"),
                    os.path.join(CSRC, "lapgpu_plan.cpp")], check=True)
    return exe


def _star(m):
    import scipy.sparse as sp

    a = sp.coo_matrix((np.ones(m), (np.zeros(m, dtype=int), np.arange(1, m + 1))), shape=(m + 1, m + 1))
    return sp.csc_matrix(a + a.T)


@pytest.mark.parametrize("name", ["grid3", "path_matchings", "dense", "star"])
def test_slot_programs_equal_the_sweeps(checker, tmp_path, name):
    """Every slot program (context tail with and without its dense block,
    the block's own programs, the batched engine's whole-graph programs),
    run lane by lane on the host as the kernels run them, equals the sweeps
    written out over the level-order layout, to rounding.  With det, rows
    split in chunks store partial sums that a fix-up step adds in chunk
    order (no atomic adds left: lead 2 count 0); partial slots start as NaN,
    so a fix-up that read an unwritten slot would fail.  The star's centre
    row (100,000 entries) needs several passes of partial slots."""
    rng = np.random.default_rng(11)
    a = {"grid3": lambda: graphs.weighted_grid3(40, 40, 40, rng, "expU2"),
         "path_matchings": lambda: graphs.path_matchings(60000, 4, rng, "logU3"),
         "dense": lambda: graphs.random_dense(1500, 120, rng),
         "star": lambda: _star(100000)}[name]()
    d = str(tmp_path / name)
    _dump(a, d)
    split_rows = 0
    for mode, dm in (("ctx", 0), ("ctx", 18432), ("wg", 0)):
        for det in ("0", "1"):
            out = subprocess.run([checker, d, mode, str(dm), det], capture_output=True, text=True,
                                 check=True).stdout.split()
            f = dict(zip(out[0::2], out[1::2]))
            for k in ("err_fwd", "err_bwd", "err_dense"):
                assert float(f[k]) < 1e-13, (mode, dm, det, k, f[k])
            npart = [int(f["npart"]), int(f["npart_dense"])]
            if det == "1":
                assert int(f["lead2"]) == 0 and max(npart) <= 1024, (mode, dm, out)
                split_rows += int(f["lead3"])
            else:
                assert int(f["lead3"]) == 0 and max(npart) == 0
    assert split_rows > 0, "expected rows split in chunks"
