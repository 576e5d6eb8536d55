"""Build of the in-tree HIP library (gfx950).

``python -m scratch_wadge_amd.build`` compiles csrc/wadges.hip into
scratch_wadge_amd/libwadges.so with hipcc.  hipcc cross-compiles, so this runs
on a machine without a GPU; the .so travels with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwadges.so")
ARCH = os.environ.get("WADGES_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions in any generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-adds are fused within a source expression only (a * b + c), never
    # across statements by the backend: the fusion then does not depend on the
    # code a routine is inlined into, so every kernel that runs the fused stage
    # body (single-problem, ensemble flat and persistent) computes the same bits
    "-ffp-contract=on",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the WADGES HIP library cannot be built")


def build(verbose=False):
    src = os.path.join(CSRC, "wadges.hip")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the disabled GPU feature
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
