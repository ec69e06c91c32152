"""Device-resident strict-pivot Cholesky against the checker driver.

The checker (oracle/cholesky_driver.py, a restatement of
python/giao_integrals/cholesky.py:60-297) runs on the C oracle's diagonal
and columns. Pivots must agree except where the two candidates' residuals
tie to 1e-12 relative (a rounding-level tie may reorder them); factors must
agree to 1e-10 up to that point; the final residual must be below the
stopping tolerance.
"""

import numpy as np
import pytest

from giaoint_amd import _native, workloads
from oracle import cholesky_driver
from oracle.pyoracle import make_basis

pytestmark = pytest.mark.gpu


def first_divergence(a, b):
    for k, (x, y) in enumerate(zip(a, b)):
        if tuple(x) != tuple(y):
            return k
    return None


@pytest.mark.parametrize("natoms", [2])
def test_device_cholesky_matches_checker(gpu, port, natoms):
    sh, B, O = workloads.cholesky_atoms(natoms=natoms)
    b = make_basis(sh.shell_list(), port)
    n = b.nao
    f = _native.field6(B, O)
    hist = []
    vec, piv, maxres, tol = cholesky_driver.cholesky_eri(
        b.ao_offsets, [((l + 1) * (l + 2) // 2, 1, 1) for l in b.L],
        lambda pairs: port.cholesky_diagonal(b, pairs, f, 16),
        lambda pv: port.cholesky_columns(b, pv, f, 16),
        real=False, single_center=False, threshold=1e-8, history=hist,
    )
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    res = _native.cholesky_run(db, f, 1e-8)
    assert res.converged
    assert res.max_residual <= res.tolerance
    assert abs(res.tolerance - tol) <= 1e-20 + 1e-12 * tol
    gp = res.pivots()
    k = first_divergence(gp, piv)
    upto = min(len(gp), len(piv)) if k is None else k
    print(f"rank gpu {res.rank} checker {len(piv)}; identical pivots for the first {upto}")
    assert upto >= min(50, len(piv))
    gv = res.vectors(0, upto)
    assert np.max(np.abs(gv - vec[:upto])) <= 1e-10
    if k is not None:
        r = hist[k]
        assert abs(r[gp[k][0] * n + gp[k][1]] - r[piv[k][0] * n + piv[k][1]]) <= 1e-12 * r.max()
    assert abs(res.rank - len(piv)) <= max(2, len(piv) // 100)


def test_device_cholesky_rank_cap_and_prefix(gpu, port):
    sh, B, O = workloads.cholesky_atoms(natoms=2)
    b = make_basis(sh.shell_list(), port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    f = _native.field6(B, O)
    with pytest.raises(MemoryError):
        _native.cholesky_run(db, f, 1e-8, max_rank=5)
    pre = _native.cholesky_run(db, f, 1e-8, stop_after=7)
    assert pre.rank == 7 and not pre.converged
