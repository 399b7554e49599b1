"""libvfrb's host code (parameter derivation, delay table, VDIF header codec,
frame sequencer, header writers) built with AddressSanitizer and
UndefinedBehaviorSanitizer and run over valid, edge-case and randomly damaged
frame streams (tests/sanitize/host_checks.cpp).  The sequencer parses
untrusted input (files, pipes): a read past a frame, an overflow in its gap
arithmetic or a leak is a sanitizer report and fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_under_asan_and_ubsan(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "host_checks"
    srcs = [os.path.join(ROOT, "vlite-frb_amd", "host", f) for f in ("vdif.cpp", "params.cpp", "writers.cpp")]
    srcs.append(os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"))
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), *srcs, "-o", str(exe), "-lpthread"]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report[-4000:]
    assert "runtime error" not in report and "Sanitizer" not in report, report[-4000:]
    assert "host checks: 0 failure(s)" in r.stdout
