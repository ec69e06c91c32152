"""Shared fixtures. Tests marked ``gpu`` need an MI355X; the rest run on CPU."""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950) and the built HIP library")


def _ensure_oracle():
    lib = ROOT / "oracle" / "liboracle.so"
    src = ROOT / "oracle" / "giao_oracle.c"
    if not lib.exists() or lib.stat().st_mtime < src.stat().st_mtime:
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "oracle"], check=True)


@pytest.fixture(scope="session")
def port():
    """The C restatement of the reference path (oracle/giao_oracle.c)."""
    _ensure_oracle()
    from oracle import pyoracle

    return pyoracle.load("port")


@pytest.fixture(scope="session")
def reference():
    """The reference's own sources compiled into oracle/_ref (skips when absent)."""
    from oracle import pyoracle

    if not pyoracle.available("reference"):
        pytest.skip("oracle/_ref/libgiao_ref.so is not built (needs the reference tree)")
    return pyoracle.load("reference")


@pytest.fixture(scope="session")
def gpu():
    """A HIP context on device 0 (GPU tests only)."""
    from giaoint_amd import _native

    return _native.Context(0)


def random_shells(rng, ls, nprim=(1, 4), ncontr=(1, 3), box=1.5, log_exp=(-0.7, 1.7)):
    out = []
    for l in ls:
        k = int(rng.integers(nprim[0], nprim[1]))
        x = int(rng.integers(ncontr[0], ncontr[1]))
        out.append((rng.uniform(-box, box, 3), int(l), 10 ** rng.uniform(*log_exp, k), rng.normal(size=(x, k))))
    return out


@pytest.fixture
def rng():
    return np.random.default_rng(20261203)
