"""The library's host code (plan, solver, C ABI) built with AddressSanitizer and
UndefinedBehaviorSanitizer against a host-only HIP stand-in, driven through every entry
point (tests/host_asan/harness.cpp).  CPU only: no kernel runs; the stand-in launchers
check every device buffer extent the real kernels touch."""
import os
import shutil
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "host_asan"))

pytestmark = pytest.mark.skipif(not (shutil.which("hipcc") or os.path.exists("/opt/rocm/bin/hipcc")),
                                reason="needs hipcc for the host-only build")


@pytest.fixture(scope="module")
def harness():
    import build_and_run
    return build_and_run.build()


def _run(exe, extra=""):
    import subprocess
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1:verify_asan_link_order=0" + (":" + extra if extra else "")
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    return subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)


def test_host_code_clean_under_asan(harness):
    res = _run(harness)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-6000:]
    assert "ALL HOST CHECKS PASSED" in res.stdout


def test_host_code_clean_with_immediate_address_reuse(harness):
    """No quarantine: freed plans and buffers are handed out again at once, which is when
    a cache keyed on an address (the solver's captured graphs) would go stale."""
    res = _run(harness, "quarantine_size_mb=0")
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-6000:]
    assert "ALL HOST CHECKS PASSED" in res.stdout
