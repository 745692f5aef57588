"""Build libodtbrain_hip.so in-tree with hipcc for gfx950.

Usage: ``python -m odtbrain_amd._build`` (or ``__graft_entry__.build()``).
"""
import os
import pathlib
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
LIB = HERE / "libodtbrain_hip.so"

SOURCES = ["odt_kernels.hip", "odt_adjoint.hip", "odt_translate.hip",
           "odt_prepare.hip", "odt_apple.hip", "odt_plan.cpp"]

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; ROCm is required to build")
    return exe


def build(verbose=False, jobs=None):
    """Compile every source to an object (in parallel) and link the shared
    library."""
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = objdir / (src + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, "-I", str(INCLUDE), "-I", str(CSRC),
               "-c", str(CSRC / src), "-o", str(obj)]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        cmds.append(cmd)
        objs.append(str(obj))

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    jobs = jobs or min(len(cmds), os.cpu_count() or 1, 8)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for f in [pool.submit(compile_one, c) for c in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs,
           "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
