"""Builds the compat parity test program (test infrastructure: it links the
CPU oracle).  Called by __graft_entry__.build()."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "compat_parity")


def build():
    lib = os.path.join(ROOT, "spm-iqc_amd", "lib")
    orc = os.path.join(ROOT, "oracle")
    cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-o", BIN, os.path.join(HERE, "compat_parity.cpp"),
           "-L", lib, "-lspm_iqc_compat", "-lspm_iqc", "-L", orc, "-llp_oracle",
           "-Wl,-rpath," + lib, "-Wl,-rpath," + orc, "-Wl,-rpath,$ORIGIN/../../spm-iqc_amd/lib",
           "-Wl,-rpath,$ORIGIN/../../oracle"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-4000:])
    return BIN


if __name__ == "__main__":
    print(build())
