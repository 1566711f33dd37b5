"""Build recipe of libswe2d_hip.so: hipcc straight to a shared object for gfx950.

Flags that matter for results:
  -ffp-contract=off   no fused multiply-add, so FP32 rounds like the reference's
                      eager ATen ops (one rounding per operation)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions (v_pk_*_f32)
HIP's default correctly rounded FP32 division and square root stay on.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SWE_OFFLOAD_ARCH", "gfx950")

SOURCES = ["csrc/swe_step.hip", "csrc/swe_adjoint.hip", "csrc/swe_ensemble.hip", "csrc/surrogate.hip",
           "csrc/probe.hip"]
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         f"--offload-arch={ARCH}", f"-I{ROOT / 'include'}", f"-I{HERE / 'csrc'}"]


def build(verbose: bool = False) -> Path:
    out = HERE / "libswe2d_hip.so"
    tmp = out.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, *[str(HERE / s) for s in SOURCES], "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}) building {out.name}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
