"""Host code of the library (settings JSON, expression parser and compiled
programs, host interface weights, bulk VTI writer) built with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer and driven over valid and
malformed inputs (tests/sanitizer/host_driver.cpp).  No GPU: the device code
is not part of this build.  The expression parser's nesting bound comes from
this test (5000 nested parentheses overflowed its stack)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nemaktis_amd", "csrc")


@pytest.mark.slow
def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ is not available")
    exe = str(tmp_path / "host_driver")
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=undefined", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, os.path.join(ROOT, "tests", "sanitizer", "host_driver.cpp")]
    cmd += [os.path.join(CSRC, f) for f in ("json.cpp", "expr.cpp", "settings.cpp", "vti.cpp")]
    cmd += ["-o", exe, "-lpthread"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(tmp_path / "bulk.vti")], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout, r.stdout[-3000:]
    assert "mismatches 0" in r.stdout
