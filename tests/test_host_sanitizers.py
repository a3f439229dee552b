"""The library's host code (mesh tables, element groups, node-block CSR patterns and gather maps, external forces,
footprint estimates, rules, shape functions, extrapolation, ABI refusals) under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/host_sanitize/host_check.cpp built from the library's own sources with the host
code instrumented (hipcc -Xarch_host -fsanitize=...), run on the CPU.  No GPU is used."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not available")
def test_host_code_is_clean_under_asan_and_ubsan():
    from pmsim_amd import build as B
    out = os.path.join(ROOT, "tests", "host_sanitize", "_build")
    os.makedirs(out, exist_ok=True)
    srcs = [os.path.join(ROOT, "tests", "host_sanitize", "host_check.cpp")] + [os.path.join(B.SRC, s) for s in B.SOURCES]
    flags = ["-x", "hip", "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer", "-Xclang",
             "-target-feature", "-Xclang", "-packed-fp32-ops"] + SAN
    procs, objs = [], []
    for s in srcs:
        o = os.path.join(out, os.path.basename(s) + ".o")
        objs.append(o)
        procs.append(subprocess.Popen([HIPCC] + flags + ["-c", s, "-o", o], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    for p in procs:
        log, _ = p.communicate()
        assert p.returncode == 0, log
    exe = os.path.join(out, "host_check")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950"] + SAN + ["-pthread", "-o", exe] + objs, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and "host checks: 0 failures" in r.stdout, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    shutil.rmtree(out, ignore_errors=True)
