"""Build the gfx950 shared library ``giaoint_amd/libgiaohip.so`` in-tree.

``python -m giaoint_amd.build`` (or ``__graft_entry__.build()``) compiles every
HIP source under ``giaoint_amd/csrc`` with hipcc for gfx950 into one
C-ABI shared library whose entry points are declared in ``include/giaohip.h``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
LIBRARY = PACKAGE / "libgiaohip.so"
ARCH = os.environ.get("GIAOHIP_ARCH", "gfx950")

SOURCES = ["gh_api.hip", "gh_eri.hip", "gh_eri_fast.hip", "gh_cholesky.hip"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-shared",
    # Packed FP32 instructions stay disabled in every HIP object.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-pass-failed",
    # compile-time plans of the class-specialized kernels
    "-fconstexpr-steps=200000000",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the gfx950 library cannot be built")


def build(verbose: bool = False) -> Path:
    """Compile the library; always rebuilds. Returns the library path."""
    sources = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{ROOT / 'include'}", f"-I{CSRC}", *sources, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with status {proc.returncode}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
