"""Builds libmlcp_hip.so for gfx950 with hipcc (in-tree).

    python -m mlcp_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "mlcp_runtime.hip")
OUT = os.path.join(HERE, "libmlcp_hip.so")
ARCH = os.environ.get("MLCP_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17",
    # packed FP32 instructions stay off (they are emitted only by hipcc itself)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP64 MFMA accumulating in VGPRs: the elimination steps read and update
    # tile elements with vector instructions between MFMAs, and the AGPR form
    # put eight accumulator copies each way and a full-latency wait around
    # every MFMA
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall", "-Wno-unused-function",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + FLAGS + ["-o", OUT + ".tmp", SRC]
    p = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (p.stdout + p.stderr).splitlines()
            if "packed-fp32-ops" not in l and l.strip()]
    if p.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
