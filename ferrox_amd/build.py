"""Build libferrox_mi.so (HIP for gfx950) and the ferrox_mi command-line
driver in-tree with hipcc.  Usage: python -m ferrox_amd.build"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libferrox_mi.so")
EXE = os.path.join(HERE, "ferrox_mi")

HIP_FLAGS = [
    "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC",
    # no packed-FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
LIB_SOURCES = ["fx_context.hip", "fx_kernels.hip", "fx_mg.hip", "fx_adjoint.hip", "fx_batch.hip", "fx_inputs.cpp", "fx_plotfile.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    return out


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    tmp = LIB + ".tmp"
    out = _run([hipcc, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    out += _run([hipcc, *HIP_FLAGS, "-o", EXE, os.path.join(CSRC, "fx_main.cpp"),
                 f"-L{HERE}", "-lferrox_mi", f"-Wl,-rpath,$ORIGIN"])
    if verbose and out.strip():
        print(out)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
