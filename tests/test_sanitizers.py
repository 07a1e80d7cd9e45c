"""The oracle and the product's host-only C++ (noise estimator, TIFF I/O,
lambda search) built with AddressSanitizer and UndefinedBehaviorSanitizer
and run on small inputs and error paths (tests/sanitize_main.cpp).  No GPU
code is compiled."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_host_code_under_asan_ubsan(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "sanitize_main")
    src = [os.path.join(ROOT, "oracle", "pgsvt_oracle.cpp"), os.path.join(ROOT, "pgure-svt_amd", "csrc", "noise.cpp"),
           os.path.join(ROOT, "pgure-svt_amd", "csrc", "tiff.cpp"), os.path.join(HERE, "sanitize_main.cpp")]
    cmd = [gxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-ffp-contract=off", "-pthread", "-o", exe] + src + ["-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "sanitized host run ok" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
