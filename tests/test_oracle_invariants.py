"""The oracle's shell drivers against closed forms and exact symmetries.

ssss closed form, pair exchange, conjugate double reversal, the one-pair
swap that is NOT a finite-field symmetry (tests/cpp/test_overlap.cpp:340-381),
contracted = sum of primitives (tests/python/test_eri_api.py:129-173), zero
field real path = general path with a vanishing field
(tests/cpp/test_overlap.cpp:575-613), and the Cholesky diagonal/columns as
slices of the full tensor.
"""

import numpy as np

from oracle.pyoracle import BasisArrays, make_basis
from tests.conftest import random_shells

FIELD = (0.3, -0.2, 0.5)
ORIGIN = (0.2, -0.1, 0.4)


def _pair(a, ea, b, eb, field, origin):
    p = ea + eb
    P = (ea * a + eb * b) / p
    ka = 0.5 * np.cross(field, a - origin)
    kb = 0.5 * np.cross(field, b - origin)
    q = kb - ka
    K = np.exp(-ea * eb / p * np.sum((a - b) ** 2))
    lon = np.exp(-q @ q / (4 * p) - 1j * (q @ P))
    return p, P - 0.5j * q / p, K * lon


def test_ssss_closed_form(port):
    A, B, C, D = (np.array(x) for x in ([-0.4, 0.2, 0.8], [0.8, -0.5, 0.1], [-0.2, 0.6, -0.3], [0.5, 0.1, -0.7]))
    e = [0.7, 1.3, 0.9, 1.1]
    f = np.array(FIELD)
    o = np.array(ORIGIN)
    p, Pc, pab = _pair(A, e[0], B, e[1], f, o)
    q, Qc, pcd = _pair(C, e[2], D, e[3], f, o)
    T = p * q / (p + q) * np.sum((Pc - Qc) ** 2)
    F0, _ = port.boys(complex(T), 0)
    norms = [(2 * x / np.pi) ** 0.75 for x in e]
    expected = 2 * np.pi**2.5 / (p * q * np.sqrt(p + q)) * pab * pcd * np.prod(norms) * F0[0]
    prims = [dict(exponent=x, center=c, angular=(0, 0, 0)) for x, c in zip(e, (A, B, C, D))]
    got = port.primitive_eri(prims, port.field6(FIELD, ORIGIN))
    assert abs(got - expected) <= 3e-13 * abs(expected)


def _higher():
    return [
        dict(exponent=0.8, center=(-0.3, 0.4, 0.1), angular=(2, 1, 0)),
        dict(exponent=1.2, center=(0.5, -0.2, 0.7), angular=(0, 1, 1)),
        dict(exponent=0.6, center=(0.2, 0.8, -0.4), angular=(1, 0, 1)),
        dict(exponent=1.4, center=(-0.7, 0.1, 0.3), angular=(1, 1, 0)),
    ]


def test_pair_symmetries(port):
    a, b, c, d = _higher()
    f = port.field6(FIELD, ORIGIN)
    v = port.primitive_eri([a, b, c, d], f)
    assert abs(v - port.primitive_eri([c, d, a, b], f)) < 8e-12
    assert abs(v - np.conj(port.primitive_eri([b, a, d, c], f))) < 8e-12
    assert abs(v - port.primitive_eri([b, a, c, d], f)) > 1e-8


def test_tiny_field_matches_real_path(port):
    rng = np.random.default_rng(9002)
    for _ in range(16):
        prims = [
            dict(exponent=float(10 ** rng.uniform(-0.4, 0.4)), center=rng.uniform(-0.8, 0.8, 3),
                 angular=tuple(int(x) for x in rng.integers(0, 2, 3)))
            for _ in range(4)
        ]
        zero = port.primitive_eri(prims, port.field6((0, 0, 0)))
        tiny = port.primitive_eri(prims, port.field6((1e-300, 0, 0)))
        assert abs(zero - tiny) <= 1e-14 + 1e-13 * abs(zero)


def test_contracted_equals_sum_of_primitives(port, rng):
    shells = random_shells(rng, [1, 0, 2, 1], nprim=(2, 4), ncontr=(1, 2))
    basis = make_basis(shells, port)
    f = port.field6(FIELD, ORIGIN)
    block, _ = port.eri_batch(basis, [(0, 1, 2, 3)], f)
    total = 0
    for ia in range(basis.nprim[0]):
        for ib in range(basis.nprim[1]):
            for ic in range(basis.nprim[2]):
                for id_ in range(basis.nprim[3]):
                    idx = (ia, ib, ic, id_)
                    sub = []
                    for s, i in enumerate(idx):
                        c, l, e, k = basis.shells[s]
                        sub.append((c, l, e[i : i + 1], k[:, i : i + 1]))
                    total = total + port.eri_batch(BasisArrays(sub), [(0, 1, 2, 3)], f)[0]
    np.testing.assert_allclose(block, total, atol=1e-13, rtol=1e-12)


def test_tensor_symmetries_and_cholesky_slices(port, rng):
    shells = random_shells(rng, [0, 1, 2, 1])
    basis = make_basis(shells, port)
    f = port.field6(FIELD, ORIGIN)
    T = port.eri_tensor(basis, f, threads=4)
    np.testing.assert_allclose(T, T.transpose(2, 3, 0, 1), atol=1e-13)
    np.testing.assert_allclose(T, T.transpose(1, 0, 3, 2).conj(), atol=1e-13)
    n = basis.nao
    pairs = [(a, b) for a in range(4) for b in range(a, 4)]
    D = port.cholesky_diagonal(basis, pairs, f, threads=4)
    mu, nu = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    np.testing.assert_allclose(D, T[mu, nu, nu, mu].real, atol=1e-13)
    off = basis.ao_offsets
    piv = [(int(off[2]) + 1, int(off[1]))]
    col = port.cholesky_columns(basis, piv, f, threads=2)[0].reshape(n, n)
    np.testing.assert_allclose(col, T[:, :, piv[0][1], piv[0][0]], atol=1e-13)


def test_schwarz_bounds_bound_every_block(port, rng):
    """|(ab|cd)| <= Q_ab Q_cd elementwise (the Cauchy-Schwarz inequality for
    the Coulomb form), Q symmetric, and Q_ab^2 at least the largest
    |(a_i b_j|b_j a_i)| (eri.cpp:587-620; test_eri_api.py:176-185)."""
    from tests.conftest import random_shells

    shells = random_shells(rng, [0, 1, 2, 1, 0])
    basis = make_basis(shells, port)
    f = port.field6((0.1, -0.3, 0.25), (0.2, 0.1, -0.1))
    Q = port.schwarz_bounds(basis, f)
    ns = len(shells)
    assert np.array_equal(Q, Q.T)
    qs = [(a, b, c, d) for a in range(ns) for b in range(ns) for c in range(ns) for d in range(ns)]
    vals, offs = port.eri_batch(basis, qs, f, 4)
    for i, (a, b, c, d) in enumerate(qs):
        blk = np.abs(vals[offs[i]:offs[i + 1]])
        assert blk.max() <= Q[a, b] * Q[c, d]
        if c == b and d == a:
            assert Q[a, b] ** 2 >= blk.max()
