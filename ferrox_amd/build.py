"""Build libferrox_mi.so (HIP for gfx950) and the ferrox_mi command-line
driver in-tree with hipcc.  Usage: python -m ferrox_amd.build

Each source is compiled to its own object in parallel, then linked."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libferrox_mi.so")
EXE = os.path.join(HERE, "ferrox_mi")
OBJ = os.path.join(HERE, "_build")

HIP_FLAGS = [
    "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC",
    # no packed-FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
LIB_SOURCES = ["fx_context.hip", "fx_kernels.hip", "fx_mg.hip", "fx_adjoint.hip", "fx_batch.hip", "fx_fft.hip", "fx_pcg.hip", "fx_inputs.cpp", "fx_plotfile.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    return out


def build(verbose: bool = False, csrc: str = CSRC, lib: str = LIB, exe: str | None = EXE, extra=()) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(os.path.dirname(lib), "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for s in LIB_SOURCES:
        src = os.path.join(csrc, s)
        obj = os.path.join(objdir, s + ".o")
        jobs.append((obj, [hipcc, *HIP_FLAGS, *extra, "-c", "-o", obj, src]))
    with ThreadPoolExecutor(max_workers=min(len(jobs), max(1, os.cpu_count() or 1))) as ex:
        outs = list(ex.map(_run, [cmd for _, cmd in jobs]))
    out = "\n".join(outs)
    tmp = lib + ".tmp"
    out += _run([hipcc, *HIP_FLAGS, "-shared", "-o", tmp, *[obj for obj, _ in jobs]])
    os.replace(tmp, lib)
    if exe:
        out += _run([hipcc, *HIP_FLAGS, "-o", exe, os.path.join(csrc, "fx_main.cpp"),
                     f"-L{os.path.dirname(lib)}", "-l" + os.path.basename(lib)[3:-3], "-Wl,-rpath,$ORIGIN"])
    if verbose and out.strip():
        print(out)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
