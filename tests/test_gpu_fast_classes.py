"""The class-specialized kernels of every class with l <= 3 and
la + lb + lc + ld = 3..12 against the oracle (each ordered class evaluated
through the symmetry image that has a kernel), in a field and at zero field,
with general contractions; and the mixed-class batch on the batched and the
plain variant. Reference contraction: src/eri.cpp:341-368."""

import itertools

import numpy as np
import pytest

from giaoint_amd import _native, workloads
from oracle.pyoracle import make_basis
from tests.conftest import random_shells

pytestmark = pytest.mark.gpu

ATOL = 1e-11


@pytest.mark.parametrize("L", range(3, 13))
def test_fast_classes_against_oracle(gpu, port, L):
    classes = [c for c in itertools.product(range(4), repeat=4) if sum(c) == L]
    worst = 0.0
    try:
        gpu.set_variant("batched")
        for cls in classes:
            rng = np.random.default_rng(sum((i + 1) * 7 ** i * l for i, l in enumerate(cls)) + 5)
            shells = random_shells(rng, list(cls), nprim=(1, 4), ncontr=(1, 3), box=1.5, log_exp=(-0.7, 1.5))
            b = make_basis(shells, port)
            db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
            q = [(0, 1, 2, 3), (2, 3, 0, 1), (1, 0, 3, 2), (3, 2, 1, 0)]
            for B in ((0.1, -0.3, 0.25), (0.0, 0.0, 0.0)):
                got = db.eri_batch(q, _native.field6(B, (0.2, 0.1, -0.1)))
                want, _ = port.eri_batch(b, q, port.field6(B, (0.2, 0.1, -0.1)), 16)
                err = float(np.max(np.abs(got - want)))
                worst = max(worst, err)
                assert err <= ATOL, f"class {cls} B={B}: max |batched - oracle| = {err:.2e}"
    finally:
        gpu.set_variant("batched")
    print(f"L={L}: {len(classes)} classes, max |batched - oracle| = {worst:.2e}")


def test_fast_classes_on_mixed_batch(gpu, port):
    """4000 mixed-class quartets (l <= 3, degree 1..8): the batched variant
    (class-specialized kernels) within 1e-11 of the plain one (the generic
    kernel, itself checked against the oracle per class)."""
    batch = workloads.mixed_batch(4000)
    sh = batch.shells
    coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
    db = _native.DeviceBasis(gpu, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)
    f = _native.field6(batch.B, batch.origin)
    try:
        gpu.set_variant("plain")
        plain = db.eri_batch(batch.quartets, f)
        gpu.set_variant("batched")
        fast = db.eri_batch(batch.quartets, f)
    finally:
        gpu.set_variant("batched")
    diff = float(np.max(np.abs(fast - plain)))
    print(f"max |batched - plain| over 4000 mixed quartets = {diff:.2e}")
    assert diff <= ATOL
