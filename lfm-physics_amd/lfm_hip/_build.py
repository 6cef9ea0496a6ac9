"""Build liblfm_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

The library is one shared object compiled from csrc/*.hip and linked against
rocFFT.  ``build()`` always recompiles; ``library_path()`` is where the
loader looks for it.
"""
from __future__ import annotations

import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)                      # lfm-physics_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB_NAME = "liblfm_hip.so"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

SOURCES = ["lfm_abi.hip", "lfm_step.hip", "lfm_scv.hip", "lfm_reduce.hip", "lfm_adjoint.hip",
           "lfm_calib.hip", "lfm_relax.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # never emit packed FP32 arithmetic (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def library_path() -> str:
    return os.path.join(PKG_DIR, LIB_NAME)


def build(verbose: bool = False, extra_flags: list[str] | None = None) -> str:
    """Compile every HIP source into ``liblfm_hip.so``; return its path."""
    out = library_path()
    tmp = out + ".tmp"
    cmd = [os.path.join(ROCM, "bin", "hipcc"), *HIPCC_FLAGS, *(extra_flags or []),
           "-I", INCLUDE, "-I", CSRC,
           *[os.path.join(CSRC, s) for s in SOURCES],
           "-L", os.path.join(ROCM, "lib"), "-lrocfft",
           f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}",
           "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr[-6000:]}")
    if verbose and r.stderr.strip():
        print(r.stderr[-3000:])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
