"""Build recipe of libsdm_hip.so (hipcc for gfx950, in-tree)."""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(PKG, "csrc", "sdm.hip")
OUT = os.path.join(PKG, "libsdm_hip.so")

# -ffp-contract=off: products and sums round as in the reference binary (x86-64 -O2 without FMA);
# -packed-fp32-ops disabled: the packed FP32 forms are avoided on this hardware.
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def build(verbose: bool = False) -> str:
    cmd = ["hipcc", *HIPCC_FLAGS, "-o", OUT, SRC]
    r = subprocess.run(cmd, cwd=os.path.dirname(SRC), capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(" ".join(cmd))
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
