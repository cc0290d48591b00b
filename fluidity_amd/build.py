"""Builds libfluidity_amd.so (the gfx950 kernels plus the C ABI) in-tree.

hipcc cross-compiles for gfx950 without a GPU.  Flags that matter:
  -ffp-contract=off      no FMA contraction: the FP64 path reproduces the
                         reference's operation order bit for bit;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                         no packed FP32 instructions;
  no -ffast-math         f64 division and sqrt stay correctly rounded.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfluidity_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["fl_sweep.hip", "fl_abi.hip", "fl_levelset.hip", "fl_ls_abi.hip", "fl_batch.hip",
           "fl_batch_abi.hip", "fl_adjoint.hip", "fl_adj_abi.hip", "fl_io.hip", "fl_noalloc.hip"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def build(verbose: bool = False, out: str | None = None, defines=()) -> str:
    """Compiles the library to `out` (default: in-tree libfluidity_amd.so);
    `defines` are extra -D macros (e.g. FL_SWEEP_WAVES=4) for tuning builds."""
    lib = out or LIB
    tag = f"{os.getpid()}"
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", f".{tag}.o"))
        cmds.append([HIPCC, *HIP_FLAGS, *[f"-D{d}" for d in defines], "-c",
                     os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    # the sources compile independently: at most `jobs` hipcc processes at once
    jobs = max(1, int(os.environ.get("MAX_JOBS") or min(8, os.cpu_count() or 1)))
    try:
        with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
            for f in [pool.submit(subprocess.check_call, cmd) for cmd in cmds]:
                f.result()
    except BaseException:
        for o in objs:
            if os.path.exists(o):
                os.remove(o)
        raise
    tmp = lib + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, lib)
    for o in objs:
        os.remove(o)
    return lib


def build_shim_driver(verbose: bool = False) -> str:
    """Compiles tests/cpp/shim_driver (a host program using the reference-named
    C++ template API) against the in-tree library."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "shim_driver.cpp")
    out = os.path.join(root, "tests", "cpp", "shim_driver")
    cmd = [HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", src, "-o", out,
           f"-L{HERE}", "-lfluidity_amd", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../fluidity_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if a != "-v"]
    out = None
    defs = []
    for a in args:
        if a.startswith("--out="):
            out = a[len("--out="):]
        elif a.startswith("-D"):
            defs.append(a[2:])
    print(build(verbose="-v" in sys.argv, out=out, defines=defs))
