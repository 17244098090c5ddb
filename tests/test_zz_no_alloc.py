"""lapgpu_batch_solve_device_noalloc against lapgpu_batch_solve_device
through the C-ABI: the same kernel on the same requests, its work vectors,
request table and graph descriptors in the library's static device scratch
instead of allocated buffers.  Two calls in a row, the second with a smaller
input (the scratch keeps the first call's contents), each compared with the
allocating entry point: the same stop reasons and iteration counts, x within
1e-12 relative (the sums are the same kernel's; only LDS atomics of rows
longer than 64 entries could reorder, and these grids have none)."""

import numpy as np
import pytest

import lapgpu
from lapgpu import graphs
from lapgpu._native import DeviceArray

pytestmark = pytest.mark.gpu


def _rhs(n, seed):
    b = np.random.default_rng(seed).standard_normal(n)
    return b - b.mean()


def _run(batch, method, gids, bs):
    db = [DeviceArray.from_host(b) for b in bs]
    dx = [DeviceArray(len(b)) for b in bs]
    infos = getattr(batch, method)(gids, [d.address for d in db], [d.address for d in dx], tol=1e-8,
                                   maxits=1000, stag_test=5)
    xs = [d.download() for d in dx]
    return xs, infos


def test_noalloc_matches_allocating_entry_point():
    gens = [graphs.weighted_grid2(60, 40, np.random.default_rng(1), "expU2"),
            graphs.weighted_grid3(14, 12, 10, np.random.default_rng(2), "expU2"),
            graphs.grid2(150, 140),  # beyond the LDS cap: y in the scratch
            graphs.weighted_grid2(9, 7, np.random.default_rng(3), "expN")]
    batch = lapgpu.Batch(0)
    try:
        gids = [batch.add_graph(lapgpu.lap(a), lapgpu.build_ldli(a, seed=1)) for a in gens]
        sizes = [a.shape[0] for a in gens]
        calls = [
            [gids[k % 4] for k in range(12)],  # first call: every graph, 12 requests
            [gids[3], gids[0], gids[3]],       # second call: smaller
        ]
        for c, ids in enumerate(calls):
            bs = [_rhs(sizes[gids.index(g)], 100 * c + k) for k, g in enumerate(ids)]
            x_ref, i_ref = _run(batch, "solve_device", ids, bs)
            x_na, i_na = _run(batch, "solve_device_noalloc", ids, bs)
            for k in range(len(ids)):
                assert i_na[k]["stop"] == i_ref[k]["stop"] == "converged", (c, k, i_na[k], i_ref[k])
                assert i_na[k]["its"] == i_ref[k]["its"], (c, k)
                assert np.linalg.norm(x_na[k] - x_ref[k]) <= 1e-12 * np.linalg.norm(x_ref[k]), (c, k)
        # beyond the scratch's maximum: refused with a message, nothing run
        big = graphs.grid2(400, 400)
        gb = batch.add_graph(lapgpu.lap(big), lapgpu.build_ldli(big, seed=1))
        ids = [gb] * 9  # 9 x 160,000 vertices x 64 B > 64 MiB
        with pytest.raises(lapgpu.LapGPUError, match="scratch"):
            _run(batch, "solve_device_noalloc", ids, [_rhs(160000, k) for k in range(9)])
    finally:
        batch.close()
