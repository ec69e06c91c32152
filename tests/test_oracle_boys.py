"""The oracle's complex Boys function against mathematics.

F_n(z) = int_0^1 t^(2n) exp(-z t^2) dt = 1F1(n+1/2; n+3/2; -z) / (2n+1)
(mpmath at 60 digits), F_n(0) = 1/(2n+1), the integration-by-parts
recurrence, conjugation symmetry and scaled/unscaled agreement; arguments
and tolerances are those of the reference's tests
(tests/python/test_boys_production.py:13-81).
"""

import cmath

import mpmath
import numpy as np
import pytest

ARGUMENTS = [
    0.0, 1.0e-14 + 1.0e-14j, 0.749999, 0.750001, 1.0 + 4.0j, 20.0 - 35.0j, -1.0 + 2.0j,
    -30.0 - 12.0j, -100.0 + 20.0j, 123.999, 124.001, 130.0 + 10.0j, 159.999j, 1000.0 - 25.0j,
]


def boys_1f1(n, z, dps=60):
    with mpmath.workdps(dps):
        zz = mpmath.mpc(z.real, z.imag)
        return complex(mpmath.hyp1f1(n + 0.5, n + 1.5, -zz) / (2 * n + 1))


def boys_quad(n, z, dps=40):
    with mpmath.workdps(dps):
        zz = mpmath.mpc(z.real, z.imag)
        return complex(mpmath.quad(lambda t: t ** (2 * n) * mpmath.exp(-zz * t * t), [0, 0.5, 1]))


def test_mpmath_routes_agree():
    for z in (1.0 + 4.0j, -12.0 + 6.0j, 0.3 - 0.2j):
        for n in (0, 5):
            assert abs(boys_1f1(n, z) - boys_quad(n, z)) <= 1e-25 + 1e-25 * abs(boys_1f1(n, z))


@pytest.mark.parametrize("argument", ARGUMENTS)
@pytest.mark.parametrize("nmax", [0, 3, 12, 32])
def test_against_hypergeometric(port, argument, nmax):
    values, _ = port.boys(argument, nmax)
    for n, v in enumerate(values):
        e = boys_1f1(n, complex(argument))
        assert abs(v - e) <= 3.0e-14 + 8.0e-13 * abs(e)


def test_zero_argument_closed_form(port):
    values, info = port.boys(0.0, 32)
    np.testing.assert_allclose(values, [1.0 / (2 * n + 1) for n in range(33)], rtol=1e-15, atol=0)
    assert info["region"] == 0


@pytest.mark.parametrize("argument", [0.2 + 0.4j, 3.0 - 7.0j, -12.0 + 6.0j])
def test_recurrence_and_conjugation(port, argument):
    v, _ = port.boys(argument, 20)
    c, _ = port.boys(argument.conjugate(), 20)
    np.testing.assert_allclose(c, v.conj(), atol=2e-14, rtol=4e-13)
    ex = cmath.exp(-argument)
    for n in range(20):
        res = (2 * n + 1) * v[n] - ex - 2 * argument * v[n + 1]
        assert abs(res) <= 3.0e-13 * max(1.0, abs((2 * n + 1) * v[n]) + abs(ex))


@pytest.mark.parametrize("argument", [-0.3 + 0.2j, -20.0 + 5.0j, -120.0 - 8.0j])
def test_scaled_unscaled(port, argument):
    u, _ = port.boys(argument, 16)
    s, _ = port.boys(argument, 16, scaled=True)
    np.testing.assert_allclose(s, cmath.exp(argument) * u, atol=2e-14, rtol=8e-13)


def test_regions(port):
    assert port.boys(0.2 + 0.1j, 4)[1]["region"] == 0
    assert port.boys(3.0 + 2.0j, 4)[1]["region"] == 1
    assert port.boys(-3.0 + 2.0j, 4)[1]["region"] == 2
    assert port.boys(1000.0 + 2.0j, 4)[1]["region"] == 3
    assert port.boys(3.0 + 2.0j, 4)[1]["segments"] >= 1


def test_errors(port):
    from oracle.pyoracle import CheckerError

    with pytest.raises(CheckerError, match="outside the verified"):
        port.boys(10.0 + 500.0j, 4)
    with pytest.raises(CheckerError, match="outside the verified"):
        port.boys(160.001j, 4)
    with pytest.raises(CheckerError, match="through 32"):
        port.boys(1.0, 33)


def test_reachable_domain(port):
    """120 seeded physical arguments (exponents 0.05..20, 4-bohr boxes,
    |B|<=1), orders up to 32 (test_boys_reachable_domain.py:8-44)."""
    rng = np.random.default_rng(410404)
    for _ in range(40):
        ea, eb = 10 ** rng.uniform(-1.3, 1.3, 2)
        p = ea + eb
        A, Bc, C = rng.uniform(-2, 2, (3, 3))
        d = rng.normal(size=3)
        field = rng.random() * d / np.linalg.norm(d)
        sep = Bc - A
        q = 0.5 * np.cross(field, sep)
        P = (ea * A + eb * Bc) / p
        Pc = P - 0.5j * q / p
        z = complex(p * np.sum((Pc - C) ** 2))
        v, _ = port.boys(z, 32)
        for n in (0, 1, 8, 16, 32):
            e = boys_1f1(n, z)
            assert abs(v[n] - e) <= 3e-14 + 8e-13 * abs(e)
