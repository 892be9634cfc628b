"""Build the gfx950 HIP library ``libnyxpm.so`` in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "nyxpm.hip")
LIB = os.path.join(HERE, "libnyxpm.so")
DEPS = [os.path.join(HERE, "csrc", f) for f in
        ("nyxpm.hip", "mg_kernels.hip", "particle_kernels.hip", "nyxpm_internal.h")] + \
       [os.path.join(ROOT, "include", "nyxpm.h")]

# -ffp-contract=off: no FMA contraction, so the FP64 path's arithmetic order is
# exactly the one written (and the one the CPU restatement uses).
# -packed-fp32-ops is switched off: the compiler must not emit packed FP32 ops.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Wno-unused-value", "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


def build(force: bool = True, verbose: bool = False) -> str:
    """Compile libnyxpm.so (always when force, else only if stale)."""
    if not force and os.path.exists(LIB):
        t = os.path.getmtime(LIB)
        if all(os.path.getmtime(d) <= t for d in DEPS):
            return LIB
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-6000:])
    if verbose:
        print(" ".join(cmd))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
