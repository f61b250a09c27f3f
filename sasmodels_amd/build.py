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


def _run(cmds):
    """Run compiler commands in parallel; (returncode, log) of the first failure or (0, logs)."""
    procs = [subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True) for c in cmds]
    logs, rc = [], 0
    for p in procs:
        out, _ = p.communicate()
        logs.append("\n".join(line for line in out.splitlines() if "packed-fp32-ops" not in line))
        rc = rc or p.returncode
    return rc, "\n".join(l for l in logs if l.strip())


def build_library(verbose=False):
    """Compile libsaship.so (one object per source, compiled in parallel, then
    linked); raises CalledProcessError with the compiler log on failure."""
    tmp = OUTPUT + ".tmp"
    objs = [os.path.join(HERE, "csrc", os.path.basename(src) + ".o") for src in SOURCES]
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    cmds = [[hipcc()] + compile_flags + ["-c", "-o", o, src] for src, o in zip(SOURCES, objs)]
    rc, log = _run(cmds)
    if rc == 0:
        link = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
        rc, link_log = _run([link])
        log = "\n".join(l for l in (log, link_log) if l.strip())
    for o in objs:
        if os.path.exists(o):
            os.remove(o)
    if rc != 0:
        sys.stderr.write(log + "\n")
        raise subprocess.CalledProcessError(rc, cmds[0], output=log)
    if verbose and log.strip():
        print(log)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
