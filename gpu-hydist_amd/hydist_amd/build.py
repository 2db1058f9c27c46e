"""In-tree build of libhydist.so for gfx950 with hipcc.

Flags: -ffp-contract=off -fno-fast-math keep every FP64 expression in the
IEEE operation order of the reference formulas (no fused multiply-add), and
the packed-FP32 target feature is disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
OUT = os.path.join(HERE, "_build", "libhydist.so")
SOURCES = ("hd_api.hip", "hd_kernels.hip", "hd_sediment.hip", "hd_cnn.hip")
ARCH = os.environ.get("HYDIST_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-function",
]


def build(verbose: bool = False) -> str:
    """Compile every HIP source into hydist_amd/_build/libhydist.so."""
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = ["hipcc"] + FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
