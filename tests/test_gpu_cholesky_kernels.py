"""GPU Cholesky diagonal and pivot columns against the oracle
(cholesky.cpp:203-314 semantics), finite and zero field."""

import numpy as np
import pytest

from giaoint_amd import _native, workloads
from oracle.pyoracle import make_basis
from tests.conftest import random_shells

pytestmark = pytest.mark.gpu


def test_diagonal_and_columns_contracted(gpu, port):
    rng = np.random.default_rng(21)
    b = make_basis(random_shells(rng, [0, 1, 2, 1, 0, 3]), port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    ns = len(b.L)
    pairs = [(x, y) for x in range(ns) for y in range(x, ns)]
    off = b.ao_offsets
    for field in (((0.2, -0.1, 0.3), (0.1, 0.0, -0.2)), ((0, 0, 0), (0, 0, 0))):
        f = _native.field6(*field)
        np.testing.assert_allclose(db.cholesky_diagonal(pairs, f), port.cholesky_diagonal(b, pairs, f, 8), atol=1e-11, rtol=0)
        piv = [(int(off[2]) + 1, int(off[5]) + 3), (int(off[2]), int(off[5]))]
        np.testing.assert_allclose(db.cholesky_columns(piv, f), port.cholesky_columns(b, piv, f, 8), atol=1e-11, rtol=0)
    z = _native.field6()
    np.testing.assert_allclose(db.cholesky_columns(piv, z, real=True), port.cholesky_columns(b, piv, z, 8, real=True),
                               atol=1e-11, rtol=0)
    with pytest.raises(ValueError, match="share a shell pair"):
        db.cholesky_columns([(0, 0), (int(off[1]), 0)], f)
    with pytest.raises(IndexError, match="ordered"):
        db.cholesky_diagonal([(1, 0)], f)


def test_cholesky_atoms_columns_sample(gpu, port):
    sh, B, O = workloads.cholesky_atoms(natoms=4)
    b = make_basis(sh.shell_list(), port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    f = _native.field6(B, O)
    ns = len(b.L)
    pairs = [(x, y) for x in range(ns) for y in range(x, ns)]
    D = db.cholesky_diagonal(pairs, f)
    np.testing.assert_allclose(D, port.cholesky_diagonal(b, pairs, f, 16), atol=1e-11, rtol=0)
    p = int(np.argmax(D))
    mu, nu = divmod(p, b.nao)
    np.testing.assert_allclose(db.cholesky_columns([(mu, nu)], f), port.cholesky_columns(b, [(mu, nu)], f, 16),
                               atol=1e-11, rtol=0)
