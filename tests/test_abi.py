"""The C ABI library without a GPU: it loads, exports every function that
include/giaohip.h declares, and its host-side shell code (contraction
normalization, types.cpp:239-306) matches the oracle."""

import re
from pathlib import Path

import numpy as np
import pytest

from giaoint_amd import _native

HEADER = Path(__file__).resolve().parents[1] / "include" / "giaohip.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(giaohip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.load_library()
    names = declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.EXPORTED)


def _port_norm(port, e, ax, ay, az):
    import ctypes as C

    fn = port.lib.or_primitive_normalization
    fn.restype = C.c_double
    fn.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int]
    return fn(e, ax, ay, az)


def test_normalization_matches_oracle(port, rng):
    for l in range(6):
        e = 10 ** rng.uniform(-1, 4, 4)
        c = rng.normal(size=(3, 4))
        ours = _native.normalize_shells([l], [4], [3], e, c.reshape(-1)).reshape(3, 4)
        np.testing.assert_array_equal(ours, port.normalize_shell(l, e, c))
        assert _native.primitive_normalization(e[0], (l, 0, 1)) == _port_norm(port, e[0], l, 0, 1)


def test_normalization_errors():
    with pytest.raises(ValueError, match="all-zero"):
        _native.normalize_shells([1], [2], [1], [1.0, 2.0], [0.0, 0.0])
    with pytest.raises(ValueError, match="finite and positive"):
        _native.normalize_shells([1], [1], [1], [-1.0], [1.0])


def test_no_gpu_raises_loudly():
    import subprocess
    import sys

    probe = (
        "import torch\n"
        "if torch.cuda.is_available():\n"
        "    print('HAS_GPU')\n"
        "else:\n"
        "    from giaoint_amd import _native\n"
        "    try:\n"
        "        _native.Context(0)\n"
        "        print('NO_RAISE')\n"
        "    except _native.GpuUnavailableError:\n"
        "        print('RAISED')\n"
    )
    out = subprocess.run([sys.executable, "-c", probe], capture_output=True, text=True, cwd=HEADER.parents[1])
    assert out.stdout.strip() in ("RAISED", "HAS_GPU"), out.stdout + out.stderr
