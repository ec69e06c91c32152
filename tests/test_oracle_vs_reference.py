"""The oracle against GIAOINT's own sources compiled into oracle/_ref.

oracle/Makefile builds types.cpp, overlap.cpp, hermite.cpp, boys.cpp,
eri.cpp and cholesky.cpp unchanged; the restatement follows their operation
order, so agreement is expected to the last bit.
"""

import numpy as np

from oracle.pyoracle import make_basis
from tests.conftest import random_shells

F = ((0.1, -0.3, 0.25), (0.2, 0.1, -0.1))


def test_boys_bitwise(port, reference):
    for z in (0.3 + 0.2j, 5.0 - 3.0j, -40.0 + 7.0j, 90.0 + 3.0j, 150j):
        for n in (0, 7, 20):
            a, ia = port.boys(z, n)
            b, ib = reference.boys(z, n)
            assert np.array_equal(a, b) and ia == ib


def test_normalization_bitwise(port, reference, rng):
    for l in range(5):
        e = 10 ** rng.uniform(-1, 3, 5)
        c = rng.normal(size=(2, 5))
        assert np.array_equal(port.normalize_shell(l, e, c), reference.normalize_shell(l, e, c))


def test_blocks_tensor_cholesky_bitwise(port, reference, rng):
    shells = random_shells(rng, [0, 1, 2, 3, 1])
    basis = make_basis(shells, port)
    f = port.field6(*F)
    qs = [(0, 1, 2, 3), (4, 3, 1, 0), (2, 2, 2, 2), (3, 3, 3, 3)]
    assert np.array_equal(port.eri_batch(basis, qs, f, 4)[0], reference.eri_batch(basis, qs, f, 4)[0])
    small = make_basis(shells[:3], port)
    assert np.array_equal(port.eri_tensor(small, f, 4), reference.eri_tensor(small, f, 4))
    pairs = [(a, b) for a in range(3) for b in range(a, 3)]
    assert np.array_equal(port.cholesky_diagonal(small, pairs, f, 4), reference.cholesky_diagonal(small, pairs, f, 4))
    off = small.ao_offsets
    piv = [(int(off[1]), int(off[2])), (int(off[1]) + 1, int(off[2]) + 2)]
    assert np.array_equal(port.cholesky_columns(small, piv, f, 4), reference.cholesky_columns(small, piv, f, 4))
    z = port.field6()
    assert np.array_equal(
        port.cholesky_columns(small, piv, z, 4, real=True), reference.cholesky_columns(small, piv, z, 4, real=True)
    )


def test_schwarz_bounds_bitwise(port, reference, rng):
    shells = random_shells(rng, [0, 1, 2, 3, 1, 0])
    basis = make_basis(shells, port)
    for f in (port.field6(*F), port.field6()):
        assert np.array_equal(port.schwarz_bounds(basis, f), reference.schwarz_bounds(basis, f))
