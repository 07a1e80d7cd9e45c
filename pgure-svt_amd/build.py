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
# gfx950 only; packed FP32 instructions disabled for all device code.  The
# run-time window instances keep the table instances' unroll pragmas on loops
# whose trip count is known at run time only: their "loop not unrolled"
# notes are silenced.
DEVICE = ["-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-Wno-pass-failed"]
HOST = ["-x", "c++", "-pthread"]

SOURCES = [
    ("kernels.hip", DEVICE),
    ("hotpixel.hip", DEVICE),
    ("mlp.hip", DEVICE),
    ("noise_gpu.hip", DEVICE),
    ("host.cpp", DEVICE),
    ("noise.cpp", HOST),
    ("tiff.cpp", HOST),
]


def _stale(obj, src):
    """True when obj is missing or older than its source or any header."""
    if not os.path.exists(obj):
        return True
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    deps.append(os.path.join(HERE, "..", "include", "pgsvt.h"))
    return any(os.path.getmtime(d) > os.path.getmtime(obj) for d in deps)


def _compile(src, flags, incremental=False):
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    if incremental and not _stale(obj, src):
        return obj
    cmd = [HIPCC] + COMMON + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + src)
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if warn.strip():
        sys.stderr.write(warn + "\n")
    return obj


def build(jobs=4, incremental=False):
    """Compile every source (incremental: only those older than a source or
    header they may include) and link the library and the front end."""
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at " + HIPCC)
    os.makedirs(OUT_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda a: _compile(a[0], a[1], incremental), SOURCES))
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
    ap.add_argument("--incremental", action="store_true", help="recompile only out-of-date objects")
    a = ap.parse_args()
    print(build(a.jobs, a.incremental))
