"""Test setup: import paths for the package and the oracle, and the `gpu` marker.

GPU tests (marked `gpu`) call liblapgpu through its C-ABI and compare with
the CPU oracle; everything else runs on the CPU.
"""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "laplacians.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs liblapgpu kernels)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    """Creates lapgpu contexts; fails loudly if the library or GPU is missing."""
    import lapgpu

    made = []

    def make(la, ldli, dense_max=None):
        ctx = lapgpu.Context(0)
        if dense_max is not None:
            ctx.set_dense_tail(dense_max)
        ctx.upload_lap(la)
        if ldli is not None:
            ctx.upload_ldli(ldli)
        made.append(ctx)
        return ctx

    yield make
    for c in made:
        c.close()
