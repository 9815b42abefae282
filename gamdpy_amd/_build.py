"""Build libgamdpy_amd.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m gamdpy_amd._build
"""
import os
import shutil
import subprocess
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(_HERE)
SOURCES = [os.path.join(_HERE, "csrc", "gmd_core.hip")]
DEPS = [os.path.join(_HERE, "csrc", f) for f in ("gmd_device.h", "gmd_kernels.h", "gmd_adjoint.h")] + \
       [os.path.join(ROOT, "include", "gamdpy_amd.h")]
OUT = os.path.join(_HERE, "libgamdpy_amd.so")
ARCH = os.environ.get("GMD_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wno-unused-result",
    # no packed FP32 instructions are generated (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build gamdpy_amd)")


def build(verbose=False):
    """Compile every HIP source into gamdpy_amd/libgamdpy_amd.so (always recompiles)."""
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include")] + SOURCES + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
