"""Host code of the library and the CPU oracle under AddressSanitizer and UBSan.

tests/sanitize/Makefile builds the library's sources (hipcc for gfx950, sanitizers on the host
side only) and the oracle into tests/_bin/, and tests/sanitize/host_checks.cpp drives the
host-only entry points through them: the level plan and the plan builder over odd frame sizes,
a rejected model description, the host NMS on tied boxes (equal to the oracle's), the CSV
reader, and the oracle's whole detection of small frames with and without threads.  Any report
of either sanitizer aborts the driver with a non-zero status.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("make") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs make and hipcc")
def test_host_code_under_asan_ubsan():
    build = subprocess.run(["make", "-s", "-j8"], cwd=os.path.join(ROOT, "tests", "sanitize"),
                           capture_output=True, text=True, timeout=900)
    assert build.returncode == 0, build.stdout[-3000:] + build.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    run = subprocess.run([os.path.join(ROOT, "tests", "_bin", "host_checks")], cwd=ROOT,
                         env=env, capture_output=True, text=True, timeout=600)
    out = run.stdout + run.stderr
    assert run.returncode == 0, out[-4000:]
    assert "host checks: 0 failures" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out
