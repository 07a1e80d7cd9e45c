"""Build libpgsvt.so (HIP kernels for gfx950 + host orchestrator) in-tree.

    python pgure-svt_amd/build.py [--jobs N]

Objects go to pgure-svt_amd/_build/, the library to
pgure-svt_amd/pguresvt/libpgsvt.so where the Python package loads it, and
the command-line front end to pgure-svt_amd/pgure-svt.
"""

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "pguresvt", "libpgsvt.so")
CLI = os.path.join(HERE, "pgure-svt")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I" + os.path.join(HERE, "..", "include")]
# gfx950 only; packed FP32 instructions disabled for all device code.
DEVICE = ["-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
HOST = ["-x", "c++", "-pthread"]

SOURCES = [
    ("kernels.hip", DEVICE),
    ("hotpixel.hip", DEVICE),
    ("mlp.hip", DEVICE),
    ("host.cpp", DEVICE),
    ("noise.cpp", HOST),
    ("tiff.cpp", HOST),
]


def _compile(src, flags):
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + src)
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if warn.strip():
        sys.stderr.write(warn + "\n")
    return obj


def build(jobs=4):
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at " + HIPCC)
    os.makedirs(OUT_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda a: _compile(*a), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("link failed")
    shutil.move(tmp, LIB)
    # command-line front end, linked against the library next to it
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "-x", "c++", os.path.join(CSRC, "cli.cpp"), "-x", "none",
           "-o", CLI + ".tmp", "-L" + os.path.dirname(LIB), "-lpgsvt", "-Wl,-rpath,$ORIGIN/pguresvt"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("cli build failed")
    shutil.move(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    print(build(a.jobs))
