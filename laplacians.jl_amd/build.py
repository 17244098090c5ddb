"""Build the native libraries in-tree.

  lapgpu/liblapgpu.so     HIP (gfx950) solve phase + C-ABI   (include/lapgpu.h)
  lapgpu/liblapfactor.so  host LDLinv builder (C++)          (include/lapfactor.h)

Usage: python laplacians.jl_amd/build.py  (or __graft_entry__.build()).
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OUT = os.path.join(HERE, "lapgpu")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-command-line-argument",
    f"-I{INC}",
    f"-I{CSRC}",
]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ...")
    elif verbose and r.stderr.strip():
        # hipcc's host pass warns about the device-only target feature
        lines = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if lines:
            print("\n".join(lines[:20]))


def build_factor(verbose: bool = False) -> str:
    out = os.path.join(OUT, "liblapfactor.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", f"-I{INC}", "-o", out, os.path.join(CSRC, "lapfactor.cpp")], verbose)
    return out


def build_gpu(verbose: bool = False) -> str:
    out = os.path.join(OUT, "liblapgpu.so")
    objs = []
    plan_o = os.path.join(OUT, "_lapgpu_plan.o")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-c", f"-I{INC}", f"-I{CSRC}", "-o", plan_o, os.path.join(CSRC, "lapgpu_plan.cpp")], verbose)
    objs.append(plan_o)
    planio_o = os.path.join(OUT, "_lapgpu_planio.o")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-c", f"-I{INC}", f"-I{CSRC}", "-o", planio_o,
          os.path.join(CSRC, "lapgpu_planio.cpp")], verbose)
    objs.append(planio_o)
    srcs = ["lapgpu.hip", "lapgpu_batch.hip"]
    outs = [os.path.join(OUT, f"_{os.path.splitext(s)[0]}.o") for s in srcs]
    with ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        list(ex.map(lambda so: _run([HIPCC, *HIP_FLAGS, "-c", "-o", so[1], os.path.join(CSRC, so[0])], verbose),
                    zip(srcs, outs)))
    objs.extend(outs)
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out, *objs], verbose)
    for o in objs:
        os.remove(o)
    return out


def build_all(verbose: bool = False) -> list[str]:
    with ThreadPoolExecutor(max_workers=2) as ex:
        fs = [ex.submit(build_factor, verbose), ex.submit(build_gpu, verbose)]
        return [f.result() for f in fs]


if __name__ == "__main__":
    for p in build_all(verbose=True):
        print("built", p)
