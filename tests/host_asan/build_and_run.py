"""Build the library's host code with AddressSanitizer against a host-only HIP stand-in
and run the C-ABI harness (no GPU needed).

    python tests/host_asan/build_and_run.py [--keep]

The library sources (mlpcg_amd/csrc) are compiled with hipcc in host-only mode, so the
kernels exist only as launch stubs; fake_kernels.cpp replaces every launcher (linked
first, --allow-multiple-definition) with a closure that checks the buffer extents the
real kernel touches, and fake_hip.cpp stands in for the HIP runtime.  Objects go to
tests/host_asan/_build/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "mlpcg_amd", "csrc")
OUT = os.path.join(HERE, "_build")
LIB_SOURCES = ["stencil.hip", "fused.hip", "tail.hip", "reduce.hip", "pcg.hip", "lanczos.hip", "misc.hip",
               "precond.cpp", "solver.cpp", "capi.cpp"]
HARNESS_SOURCES = ["fake_kernels.cpp", "fake_hip.cpp", "harness.cpp"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


# host-only compilation: the sanitizer flag applies to the host pass only
FLAGS = ["-x", "hip", "--offload-host-only", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer",
         "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
         f"-I{os.path.join(ROOT, 'include')}", "-Wno-unused-command-line-argument"]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src) if src in LIB_SOURCES else os.path.join(HERE, src)
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    cmd = [_hipcc(), *FLAGS, "-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stderr}")
    return obj


def build() -> str:
    os.makedirs(OUT, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        lib_objs = list(ex.map(_compile, LIB_SOURCES))
        fake_objs = list(ex.map(_compile, HARNESS_SOURCES))
    # host-only objects still name the device code images they would register
    nm = subprocess.run(["nm", "-u", *lib_objs], capture_output=True, text=True, check=True).stdout
    fatbins = sorted(set(re.findall(r"\b(__hip_fatbin_[0-9a-f]+)\b", nm)))
    stub = os.path.join(OUT, "fatbin_stubs.c")
    with open(stub, "w") as f:
        for name in fatbins:
            f.write(f"const char {name}[16] = {{0}};\n")
    stub_obj = os.path.join(OUT, "fatbin_stubs.o")
    subprocess.run(["gcc", "-c", stub, "-o", stub_obj], check=True)
    exe = os.path.join(OUT, "host_harness")
    cmd = [_hipcc(), "--offload-host-only", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
           "-fsanitize=undefined", "-Wl,--allow-multiple-definition", "-o", exe, *fake_objs, *lib_objs, stub_obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stderr}")
    return exe


def run(exe: str) -> subprocess.CompletedProcess:
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1:verify_asan_link_order=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    return subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)


def main() -> int:
    exe = build()
    res = run(exe)
    sys.stdout.write(res.stdout)
    sys.stderr.write(res.stderr)
    return res.returncode


if __name__ == "__main__":
    sys.exit(main())
