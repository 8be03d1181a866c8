"""Build libmlpcg_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m mlpcg_amd.build [--jobs N] [--verbose]

Objects go to mlpcg_amd/_build/, the library to mlpcg_amd/lib/libmlpcg_hip.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmlpcg_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

SOURCES = ["stencil.hip", "fused.hip", "tail.hip", "reduce.hip", "pcg.hip", "lanczos.hip", "misc.hip",
           "precond.cpp", "solver.cpp", "capi.cpp"]

ARCH = os.environ.get("MLPCG_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    # packed FP32 instructions stay off (v_pk_* FP32 forms)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


# per-source flags: the fused apply keeps MFMA operands and results in VGPRs (no AGPR
# copies in its inner loop), and its LDS reads stay 16-bit: without unaligned-access-mode
# the compiler does not merge two neighbouring 16-bit reads at an odd element into one
# misaligned 32-bit read, which measured 5x slower on gfx950
EXTRA = {"fused.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form",
                       "-Xclang", "-target-feature", "-Xclang", "-unaligned-access-mode"]}


def _compile(src: str, verbose: bool) -> str:
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *CFLAGS, *EXTRA.get(src, []), "-c", os.path.join(SRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stderr}")
    noise = "is not a recognized feature for this target"
    errs = [ln for ln in res.stderr.splitlines() if ln.strip() and noise not in ln]
    if errs and verbose:
        print("\n".join(errs), file=sys.stderr)
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args(argv)
    print(build(args.jobs, args.verbose))


if __name__ == "__main__":
    main()
