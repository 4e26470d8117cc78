"""Build ``matvans_amd/lib/libmatvans.so`` from ``matvans_amd/csrc`` with hipcc (gfx950),
and the host driver ``matvans_amd/lib/matvans_run`` from ``matvans_amd/driver``.

    python -m matvans_amd.build [--jobs N] [--asm DIR]

Every HIP translation unit is compiled with packed-FP32 instructions disabled.  The
library is built in-tree so that it travels with the source tree.
"""
from __future__ import annotations

import argparse
import hashlib
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libmatvans.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-I", os.path.join(HERE, "..", "include")]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


STAMP = os.path.join(LIBDIR, "libmatvans.sources")
HEADER = os.path.join(HERE, "..", "include", "matvans.h")


def source_digest() -> str:
    """sha256 over the library's sources (csrc/*.hip, *.cpp, *.h and include/matvans.h),
    the stamp that ties a built libmatvans.so to the sources it was built from."""
    h = hashlib.sha256()
    files = sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp", ".h")))
    for path in [os.path.join(CSRC, f) for f in files] + [HEADER]:
        h.update(os.path.basename(path).encode() + b"\0")
        with open(path, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def _compile(src, asm_dir=None):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip"] + COMMON[:] + ["-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stderr}")
    if asm_dir and src.endswith(".hip"):
        os.makedirs(asm_dir, exist_ok=True)
        out = os.path.join(asm_dir, os.path.basename(src) + ".s")
        subprocess.run([HIPCC] + COMMON + ["--cuda-device-only", "-S", src, "-o", out],
                       check=True, capture_output=True)
    return obj, res.stderr


def build(jobs: int | None = None, asm_dir: str | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, asm_dir), srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                sys.stderr.write(err)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    with open(STAMP, "w") as fh:
        fh.write(source_digest() + "\n")
    build_driver()
    return LIB


DRIVER_SRC = os.path.join(HERE, "driver", "matvans_run.cpp")
DRIVER = os.path.join(LIBDIR, "matvans_run")


def build_driver() -> str:
    """The C++ host driver (the reference's chained benchmark harness) linked against the
    library, with an rpath to find it next to itself."""
    cmd = [HIPCC, "-x", "hip"] + COMMON + [DRIVER_SRC, "-o", DRIVER + ".tmp", "-L", LIBDIR,
                                          "-lmatvans", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{res.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--asm", default=None, help="also write device assembly to this dir")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.asm, a.verbose))


if __name__ == "__main__":
    main()
