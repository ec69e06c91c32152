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


def test_batch_error_from_either_part(gpu, port):
    """A Boys domain error in a batch that has both batched (class-specialized)
    and plain (lane-per-primitive-quartet) requests is reported whichever part
    it comes from, and the context computes the next batch correctly. The
    failing quartet's Boys argument is -225 (|z| > 160, the reference's
    verified direct domain), as the oracle reports for it too."""
    from oracle.pyoracle import CheckerError

    far = [(np.array([3.0, 0.0, 0.0]), 0, np.array([0.01]), np.array([[1.0]])),
           (np.array([-3.0, 0.0, 0.0]), 0, np.array([0.01]), np.array([[1.0]])),
           (np.array([3.0, 0.0, 0.0]), 1, np.array([0.01]), np.array([[1.0]]))]
    rng = np.random.default_rng(17)
    near = random_shells(rng, [2, 2, 1, 0], nprim=(1, 3), ncontr=(1, 2), box=1.0, log_exp=(-0.3, 1.0))
    b = make_basis(far + near, port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    B, O = (0.0, 0.0, 1.0), (0.0, 0.0, 0.0)
    f = _native.field6(B, O)
    good = [(3, 4, 3, 4), (3, 4, 4, 3), (5, 6, 5, 6), (3, 3, 5, 6), (6, 6, 6, 6), (5, 6, 6, 6)]
    with pytest.raises(CheckerError, match="outside the verified"):
        port.eri_batch(b, [(0, 1, 1, 0)], port.field6(B, O), 1)
    gpu.set_variant("batched")
    # the failing quartet in the batched part ((ss|ss) has a class kernel)
    with pytest.raises(_native.BoysNumericalError, match="outside the verified"):
        db.eri_batch(good + [(0, 1, 1, 0)], f)
    # in the plain part ((ss|sp) runs on the lane-per-primitive-quartet kernel)
    with pytest.raises(_native.BoysNumericalError, match="outside the verified"):
        db.eri_batch(good + [(0, 1, 1, 2)], f)
    got = db.eri_batch(good, f)
    want, _ = port.eri_batch(b, good, port.field6(B, O), 16)
    assert float(np.max(np.abs(got - want))) <= ATOL
