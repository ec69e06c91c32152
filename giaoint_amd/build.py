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
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
LIBRARY = PACKAGE / "libgiaohip.so"
OBJDIR = PACKAGE / "_build"
ARCH = os.environ.get("GIAOHIP_ARCH", "gfx950")

SOURCES = [
    "gh_api.hip",
    "gh_eri.hip",
    "gh_cholesky.hip",
    "gh_deriv.hip",
    "gh_eri_fast.hip",
    "gh_fast_dddd.hip",
    "gh_fast_ddff.hip",
    "gh_fast_pppp.hip",
    "gh_fast_psps.hip",
    "gh_fast_ssss.hip",
    "gh_fast_g0.hip",
    "gh_fast_g1.hip",
    "gh_fast_g2.hip",
    "gh_fast_g3.hip",
    "gh_fast_g4.hip",
    "gh_fast_g5.hip",
    "gh_fast_g6.hip",
    "gh_fast_g7.hip",
    "gh_fast_g8.hip",
    "gh_fast_g9.hip",
    "gh_fast_g10.hip",
    "gh_fast_g11.hip",
    "gh_fast_g12.hip",
    "gh_fast_g13.hip",
    "gh_fast_g14.hip",
    "gh_fast_g15.hip",
    "gh_fast_g16.hip",
    "gh_mfma_a.hip",
    "gh_mfma_b.hip",
    "gh_mfma_c.hip",
    "gh_mfma_s.hip",
    "gh_small.hip",
]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    # Packed FP32 instructions stay disabled in every HIP object.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-pass-failed",
    # compile-time plans of the class-specialized kernels
    "-fconstexpr-steps=200000000",
]


# rocBLAS: the plain complex GEMMs of the exchange contraction
LINK_LIBS = [f"-L{os.environ.get('ROCM_PATH', '/opt/rocm')}/lib", "-lrocblas",
             f"-Wl,-rpath,{os.environ.get('ROCM_PATH', '/opt/rocm')}/lib"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the gfx950 library cannot be built")


def _stale(obj: Path, src: Path) -> bool:
    if not obj.exists():
        return True
    newest = max([src.stat().st_mtime] + [h.stat().st_mtime for h in CSRC.glob("*.h")] +
                 [h.stat().st_mtime for h in (ROOT / "include").glob("*.h")])
    return newest >= obj.stat().st_mtime


def build(verbose: bool = False, jobs: int | None = None, incremental: bool = False) -> Path:
    """Compile every source to an object (in parallel), then link the library.

    Rebuilds everything unless `incremental` (then only objects older than
    their source or any header). Returns the library path.
    """
    missing = [s for s in SOURCES if not (CSRC / s).exists()]
    if missing:
        raise RuntimeError(f"missing HIP sources: {missing}")
    OBJDIR.mkdir(exist_ok=True)
    cc = hipcc()
    include = [f"-I{ROOT / 'include'}", f"-I{CSRC}"]

    def compile_one(name: str) -> Path:
        obj = OBJDIR / (Path(name).stem + ".o")
        if incremental and not _stale(obj, CSRC / name):
            return obj
        cmd = [cc, *HIPCC_FLAGS, *include, "-c", str(CSRC / name), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            sys.stderr.write(proc.stdout + proc.stderr)
            raise RuntimeError(f"hipcc failed on {name} with status {proc.returncode}")
        return obj

    workers = jobs or min(len(SOURCES), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        objects = list(pool.map(compile_one, SOURCES))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *[str(o) for o in objects], *LINK_LIBS, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc link failed with status {proc.returncode}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv))
