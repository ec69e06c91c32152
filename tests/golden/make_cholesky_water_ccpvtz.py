"""Writes cholesky_water_ccpvtz.json: the checker's pivot sequences for water
in cc-pVTZ (65 Cartesian AOs), the molecule of the reference's
tests/python/test_cholesky_molecules.py:9-14 and
benchmarks/cholesky_molecules.py:25-49, in the three runs its recorded
results cover (benchmarks/results/cholesky_molecules_0_11_1_macos_arm64.json):
zero field, the field (0.03, -0.02, 0.05) with gauge origin (0.7, -0.4, 1.1)
with strict pivoting (pivot_batch_size=1), and the same field with the
candidate-block policy (pivot_batch_size=32). Threshold 1e-8.

The checker is oracle/cholesky_driver.py (the reference driver restated)
over the selected integrals of the reference's own C++ sources compiled into
oracle/_ref (the C restatement in oracle/ gives bitwise the same integrals,
only slower). The same runs over the build of those sources with fused
multiply-add contraction (oracle/_ref/libgiao_ref_fma.so) are stored under
"<case>_fma": the candidate-block policy is sensitive to rounding at the
level of one unit in the last place, and FMA contraction alone moves its
rank from 687 to 688. Run: python tests/golden/make_cholesky_water_ccpvtz.py
"""

from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from giaoint_amd import workloads  # noqa: E402
from oracle import cholesky_driver, pyoracle  # noqa: E402

WATER_BOHR = [("O", (0.0, 0.1, 0.2)), ("H", (0.1, 1.4, -0.9)), ("H", (-0.2, -1.3, -1.0))]
FIELD = ((0.03, -0.02, 0.05), (0.7, -0.4, 1.1))
CASES = {"zero": (None, 32), "strict": (FIELD, 1), "policy": (FIELD, 32)}
OUT = Path(__file__).resolve().parent / "cholesky_water_ccpvtz.json"


def water_shells():
    data = json.loads((workloads.DATA / "cc-pvtz-h-o.json").read_text())
    return workloads.bse_shells(WATER_BOHR, data)


def run(checker, case: str, max_pivots=None, follow=None, stats=None):
    field, batch = CASES[case]
    shells = water_shells()
    basis = pyoracle.make_basis(shells, checker)
    f = checker.field6(*field) if field else checker.field6()
    real = field is None
    desc = [((l + 1) * (l + 2) // 2, int(np.atleast_2d(c).shape[0]), len(e)) for _, l, e, c in shells]
    return cholesky_driver.cholesky_eri(
        basis.ao_offsets, desc,
        lambda pairs: checker.cholesky_diagonal(basis, pairs, f, 8),
        lambda piv: checker.cholesky_columns(basis, piv, f, 8, real=real),
        real=real, single_center=False, threshold=1e-8, pivot_batch_size=batch, max_pivots=max_pivots,
        follow=follow, stats=stats,
    )


def pivots_sha256(pivots) -> str:
    return hashlib.sha256(memoryview(np.ascontiguousarray(pivots, dtype=np.int64))).hexdigest()


def main():
    out = {}
    for kind, suffix in (("reference", ""), ("reference_fma", "_fma")):
        checker = pyoracle.load(kind)
        for case in CASES:
            vectors, pivots, residual, _ = run(checker, case)
            out[case + suffix] = {"rank": int(len(pivots)), "residual": residual,
                                  "pivots_sha256": pivots_sha256(pivots), "pivots": pivots.astype(int).tolist()}
            print(kind, case, len(pivots), repr(residual), flush=True)
    OUT.write_text(json.dumps(out, separators=(",", ":")) + "\n")


if __name__ == "__main__":
    main()
