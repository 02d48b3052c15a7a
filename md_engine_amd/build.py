"""Build libmdx.so (HIP for gfx950) in-tree with hipcc.

Each translation unit under csrc/ is compiled to an object in parallel and the
objects are linked into md_engine_amd/libmdx.so together with hipFFT.  The
neighbor-list unit is compiled with -ffp-contract=off: its FP64 distance test
must round exactly like the CPU oracle's.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmdx.so")

SOURCES = ["api.hip", "nlist.hip", "pair.hip", "pppm.hip", "e3b.hip", "integrate.hip", "rigid.hip",
           "surrogate.hip", "box.hip", "flex.hip", "ring.hip"]
NO_CONTRACT = {"nlist.hip"}

ARCH = os.environ.get("MDX_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for p in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if p and os.path.exists(p):
            return p
    raise RuntimeError("hipcc not found (ROCm required to build libmdx)")


def common_flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        # packed FP32 instructions are kept out of every kernel
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        "-Wno-unused-result",
    ]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [hipcc(), *common_flags(), "-c", os.path.join(CSRC, src), "-o", out]
    if src in NO_CONTRACT:
        cmd.insert(1, "-ffp-contract=off")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
