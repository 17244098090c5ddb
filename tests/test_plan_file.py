"""The plan file (laplacians.jl_amd/csrc/lapgpu_planio.cpp), on the host: a
plan written and read back equals the one built, array for array and bit
for bit (level order, permuted Laplacian, both sweeps, tail and dense-block
programs, level groups); a file with a flipped byte, a cut end, another
magic or version is refused with a message, never half loaded."""

import os
import subprocess

import numpy as np
import pytest

from lapgpu import graphs
from test_plan_host import CSRC, ROOT, _dump


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("planio") / "plan_io_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{ROOT}/include", f"-I{CSRC}", "-o", exe,
                    os.path.join(ROOT, "tests", "native", "plan_io_check.cpp"),
                    os.path.join(CSRC, "lapgpu_plan.cpp"), os.path.join(CSRC, "lapgpu_planio.cpp")], check=True)
    return exe


def _run(*args):
    return subprocess.run(list(args), capture_output=True, text=True, check=True).stdout.split()


@pytest.mark.parametrize("name", ["grid3", "path_matchings", "dense"])
def test_plan_file_round_trip(tool, tmp_path, name):
    rng = np.random.default_rng(21)
    a = {"grid3": lambda: graphs.weighted_grid3(40, 40, 40, rng, "expU2"),
         "path_matchings": lambda: graphs.path_matchings(60000, 4, rng, "logU3"),
         "dense": lambda: graphs.random_dense(1500, 120, rng)}[name]()
    d = str(tmp_path / "in")
    _dump(a, d)
    f = str(tmp_path / "p.plan")
    out = _run(tool, "save", d, f)
    assert out[:2] == ["equal", "1"], out
    assert int(out[out.index("K") + 1]) > 0  # a dense block travels too
    assert not os.path.exists(f + ".tmp")
    assert _run(tool, "load", f) == ["ok", "n", str(a.shape[0])]


def test_plan_file_refuses_damaged_files(tool, tmp_path):
    a = graphs.weighted_grid2(80, 60, np.random.default_rng(22), "expU2")
    d = str(tmp_path / "in")
    _dump(a, d)
    f = str(tmp_path / "p.plan")
    assert _run(tool, "save", d, f)[:2] == ["equal", "1"]
    raw = open(f, "rb").read()

    def load(data, name):
        p = str(tmp_path / name)
        open(p, "wb").write(data)
        return " ".join(_run(tool, "load", p))

    # one payload byte flipped in the middle, and in the checksum itself
    mid = bytearray(raw)
    mid[len(raw) // 2] ^= 0x40
    assert "checksum" in load(bytes(mid), "flip.plan")
    last = bytearray(raw)
    last[-1] ^= 1
    assert "checksum" in load(bytes(last), "flip2.plan")
    # cut short, and one byte too many
    assert load(raw[: len(raw) - 100], "cut.plan").startswith("error")
    assert "checksum" in load(raw + b"\0", "long.plan")
    # not a plan file / another version
    assert "not a plan file" in load(b"XXXXXXXX" + raw[8:], "magic.plan")
    assert "version" in load(raw[:8] + (2).to_bytes(4, "little") + raw[12:], "version.plan")
    assert load(b"", "empty.plan").startswith("error")
