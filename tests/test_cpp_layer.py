"""The C++ host layer (include/giao_hip.hpp) without a GPU: it compiles with
g++ -std=c++20 -Wall -Wextra -Werror against the reference's entry-point
names, maps the library's shell validation to std::invalid_argument with its
messages, and raises std::runtime_error when no GPU is present."""

import numpy as np

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
