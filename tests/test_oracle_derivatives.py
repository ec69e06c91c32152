"""The oracle's ERI centre derivatives (derivatives.cpp:54-84,232-241,556-660,
1307-1415) pinned three ways:

* against GIAOINT's own derivatives.cpp compiled here (oracle/_ref), bit for
  bit, blocks and derivative Schwarz bounds;
* against Richardson-extrapolated central differences of the oracle's ERIs
  with one quartet slot's centre moved (the step and extrapolation of
  tests/python/test_moment_derivatives.py:20-23), which pins the shifted-
  angular-momentum terms and the London phase terms from the mathematics;
* against the translation sum rule: moving all four centres by d multiplies
  (ab|cd) by exp(i/2 (B x (A - B) + B x (C - D)) . d), so the four slot
  derivatives along an axis add up to i/2 (B x (A-B) + B x (C-D)) (ab|cd).

And the batch checks of compute_eri_derivative_blocks (derivatives.cpp:
1392-1408) with their messages; the derivative bounds dominate every block
(tests/python/test_eri_derivative_batches.py:59-73).
"""

from itertools import product

import numpy as np
import pytest

from oracle.pyoracle import BasisArrays, CheckerError

FIELDS = [(0.0,) * 6, (0.17, -0.11, 0.23, 0.2, -0.3, 0.1)]


def richardson(function, step=4e-4):
    coarse = (function(step) - function(-step)) / (2 * step)
    fine = (function(step / 2) - function(-step / 2)) / step
    return (4 * fine - coarse) / 3


def _shells(port):
    raw = [((0.1, -0.2, 0.3), 1, [0.7, 1.2], [[0.6, -0.2], [0.3, 0.8]]),
           ((-0.3, 0.4, 0.2), 2, [0.9], [[1.0]]),
           ((0.2, 0.5, -0.4), 0, [0.8, 1.2], [[0.7, -0.2]])]
    return [(c, l, e, port.normalize_shell(l, e, k)) for c, l, e, k in raw]


QUARTETS = [(0, 1, 2, 0), (1, 1, 0, 2), (2, 2, 2, 2), (0, 0, 1, 1), (1, 0, 1, 0), (2, 1, 1, 0)]


@pytest.mark.parametrize("field", FIELDS)
def test_port_equals_reference_build(port, reference, field):
    b = BasisArrays(_shells(port))
    f = np.array(field)
    vp, op = port.eri_derivative_blocks(b, QUARTETS, f)
    vr, orf = reference.eri_derivative_blocks(b, QUARTETS, f)
    assert np.array_equal(op, orf)
    assert np.array_equal(vp, vr)
    assert np.array_equal(port.eri_derivative_pair_bounds(b, f), reference.eri_derivative_pair_bounds(b, f))


@pytest.mark.parametrize("field", FIELDS)
def test_finite_differences_and_sum_rule(port, field):
    sh = _shells(port)
    f = np.array(field)
    for q in QUARTETS:
        four = [sh[k] for k in q]  # independent copies: each slot moves alone
        b = BasisArrays(four)
        d = port.eri_derivative_blocks(b, [(0, 1, 2, 3)], f)[0].reshape(4, 3, -1)
        value = port.eri_batch(b, [(0, 1, 2, 3)], f)[0]
        scale = max(float(np.max(np.abs(d))), float(np.max(np.abs(value))))
        for slot in range(4):
            for axis in range(3):
                def moved(step):
                    s2 = list(four)
                    c = np.array(s2[slot][0], float)
                    c[axis] += step
                    s2[slot] = (c,) + tuple(s2[slot][1:])
                    return port.eri_batch(BasisArrays(s2), [(0, 1, 2, 3)], f)[0]

                np.testing.assert_allclose(d[slot, axis], richardson(moved), atol=1e-10 * scale, rtol=0)
        A = [np.asarray(s[0], float) for s in four]
        phase = 0.5j * (np.cross(f[:3], A[0] - A[1]) + np.cross(f[:3], A[2] - A[3]))
        np.testing.assert_allclose(d.sum(axis=0), phase[:, None] * value[None, :], atol=1e-14, rtol=0)


@pytest.mark.parametrize("field", FIELDS)
def test_bounds_dominate_blocks(port, field):
    b = BasisArrays(_shells(port))
    f = np.array(field)
    D = port.eri_derivative_pair_bounds(b, f)
    Q = port.schwarz_bounds(b, f)
    qs = list(product(range(3), repeat=4))
    v, offs = port.eri_derivative_blocks(b, qs, f)
    for i, (a, bb, c, d) in enumerate(qs):
        bound = max(D[a, bb] * Q[c, d], Q[a, bb] * D[c, d])
        assert np.max(np.abs(v[offs[i]:offs[i + 1]])) <= bound * (1 + 1e-13)


def test_batch_checks_and_messages(port, reference):
    b = BasisArrays(_shells(port))
    f = np.zeros(6)
    size = 12 * b.quartet_size((0, 1, 2, 0))
    cases = [
        (np.array([1, 1 + size], dtype=np.uint64), 1 + size, 2, "derivative batch offsets have an incorrect size or start"),
        (np.array([0, size - 1], dtype=np.uint64), size - 1, 2, "derivative batch offsets do not match block sizes"),
        (np.array([0, size], dtype=np.uint64), size + 3, 2, "derivative batch output has an incorrect size"),
    ]
    for checker in (port, reference):
        for offsets, out_len, code, message in cases:
            with pytest.raises(CheckerError) as err:
                checker.eri_derivative_blocks(b, [(0, 1, 2, 0)], f, offsets=offsets, out_len=out_len)
            assert err.value.code == code and err.value.message == message
        with pytest.raises(CheckerError) as err:
            checker.eri_derivative_blocks(b, [(0, 1, 2, 5)], f, offsets=np.array([0, size], dtype=np.uint64),
                                          out_len=size)
        assert err.value.code == 3 and err.value.message == "derivative shell quartet index is out of range"
