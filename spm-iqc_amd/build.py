"""Build the gfx950 shared library libspm_iqc.so from csrc/*.hip.

Every translation unit is compiled by hipcc for gfx950 only, in parallel,
into _build/, then linked into lib/libspm_iqc.so (in-tree, so the library
travels with the source tree).  Run `python spm-iqc_amd/build.py`.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libspm_iqc.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 instructions stay off: some of their forms have given wrong results on gfx950
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-6000:]))
    os.replace(tmp, LIB)
    _build_compat()
    if verbose:
        print("built", LIB)
    return LIB


COMPAT_LIB = os.path.join(LIBDIR, "libspm_iqc_compat.so")
STAGE1 = os.path.join(LIBDIR, "spm_stage1")
STAGE2 = os.path.join(LIBDIR, "spm_stage2")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-6000:]))


def _build_compat():
    """The reference-signature C++ layer over the C ABI, and the stage-1 and stage-2 drivers."""
    src = os.path.join(HERE, "compat")
    host = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(HERE, "..", "include")]
    # the compat layer holds a few elementwise kernels of its own (BasicOperators helpers)
    dev = [HIPCC, "-x", "hip", f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-fPIC", "-Wall",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", os.path.join(HERE, "..", "include")]
    _run(dev + ["-shared", "-o", COMPAT_LIB, os.path.join(src, "compat.cpp"), "-L", LIBDIR, "-lspm_iqc",
                "-Wl,-rpath,$ORIGIN"])
    _run(host + ["-o", STAGE1, os.path.join(src, "main_spm.cpp"), "-L", LIBDIR, "-lspm_iqc_compat", "-lspm_iqc",
                 "-Wl,-rpath,$ORIGIN"])
    _run(host + ["-o", STAGE2, os.path.join(src, "main_grad.cpp"), "-L", LIBDIR, "-lspm_iqc_compat", "-lspm_iqc",
                 "-Wl,-rpath,$ORIGIN"])


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else None)
