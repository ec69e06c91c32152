"""The C++ host layer (include/giao_hip.hpp) without a GPU: it compiles with
g++ -std=c++20 -Wall -Wextra -Werror against the reference's entry-point
names, maps the library's shell validation to std::invalid_argument with its
messages, and raises std::runtime_error when no GPU is present."""

import numpy as np
import pytest

from giaoint_amd import _native, core
from tests import cpp_layer


def _has_gpu() -> bool:
    try:
        _native.Context(0)
        return True
    except Exception:
        return False


def test_host_side(tmp_path):
    exe = cpp_layer.build(tmp_path)
    r = cpp_layer.run(exe, "host", tmp_path)
    assert r["shell_count"].startswith("invalid_argument")
    assert r["shell_exponent"] == "invalid_argument shell exponents must be finite and positive"
    assert r["primitive_exponent"] == "invalid_argument Gaussian exponent must be finite and positive"
    got = tuple(int(x) for x in r["canonical"].split())
    assert got == (*core._canonical_shell_quartet(3, 1, 0, 2)[0], 0)
    assert r["shell"] == "2 2 6 12"
    kappa = np.array([float(x) for x in r["kappa"].split()])
    B, O, C = np.array([0.1, -0.2, 0.3]), np.array([1.0, 0.0, -1.0]), np.array([0.5, 2.0, -3.0])
    np.testing.assert_allclose(kappa, 0.5 * np.cross(B, C - O), rtol=0, atol=1e-16)
    expected = "ok" if _has_gpu() else "runtime_error no usable GPU for the giaohip library (device 0)"
    assert r["boys_without_gpu"] == expected
    # Nucleus (types.cpp:188-194)
    assert r["nucleus_zero_charge"] == "invalid_argument nuclear charge must be finite and positive"
    assert r["nucleus_negative_charge"] == "invalid_argument nuclear charge must be finite and positive"
    assert r["nucleus_nan_center"] == "invalid_argument nuclear center values must be finite"
    assert r["nucleus_ok"] == "ok"
    # primitive_normalization (types.cpp:60-72): N = (2a/pi)^(3/4) (4a)^(L/2) / sqrt((2ax-1)!!(2ay-1)!!(2az-1)!!)
    a = 0.7
    want = (2 * a / np.pi) ** 0.75 * (4 * a) ** 1.5 / np.sqrt(3.0 * 1.0 * 1.0)
    assert float(r["primitive_normalization"]) == pytest.approx(want, rel=1e-14)
    # the normalized primitive has unit self-overlap: int x^4 y^2 e^{-2a r^2} = 3 pi^(3/2) / (4 (2a)^(7/2) ... )
    overlap = want**2 * (3 / (4 * (2 * a) ** 2)) * (1 / (2 * (2 * a))) * (np.pi / (2 * a)) ** 1.5
    assert overlap == pytest.approx(1.0, rel=1e-14)
    assert r["normalization_exponent"] == "invalid_argument Gaussian exponent must be finite and positive"
    assert r["components"] == "10 300 210 201 120 111 102 030 021 012 003"
    assert 1 <= int(r["max_threads"]) <= 16
