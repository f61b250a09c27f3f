"""
Build libsaship.so in-tree for gfx950 with hipcc.

    python -m sasmodels_amd.build
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = [os.path.join(HERE, "csrc", name) for name in ("saship.hip", "batch.hip", "mlp.hip")]
OUTPUT = os.path.join(HERE, "libsaship.so")
ARCH = os.environ.get("SASHIP_ARCH", "gfx950")

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wall", "-Wno-unused-value", "-Wno-unused-result",
    # no packed FP32 instructions in the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    """Path of hipcc (ROCm install or PATH)."""
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose=False):
    """Compile libsaship.so; raises CalledProcessError with the compiler log on failure."""
    tmp = OUTPUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + SOURCES
    proc = subprocess.run(cmd, capture_output=True, text=True)
    log = "\n".join(line for line in (proc.stdout + proc.stderr).splitlines()
                    if "packed-fp32-ops" not in line)
    if proc.returncode != 0:
        sys.stderr.write(log + "\n")
        raise subprocess.CalledProcessError(proc.returncode, cmd, output=log)
    if verbose and log.strip():
        print(log)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
