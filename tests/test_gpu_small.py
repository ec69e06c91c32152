"""The lane-per-primitive-quartet kernel of the smallest classes
(la + lb + lc + ld <= 2, gh_small.hip) against the oracle: every class, in a
field and at zero field, general contractions, partial ket ranges through the
Cholesky column API; and the mixed-class batch with the kernel on and off.
Reference contraction: src/eri.cpp:341-368."""

import itertools

import numpy as np
import pytest

from giaoint_amd import _native, workloads
from oracle.pyoracle import make_basis
from tests.conftest import random_shells

pytestmark = pytest.mark.gpu

ATOL = 1e-11
CLASSES = [c for c in itertools.product(range(3), repeat=4) if sum(c) <= 2]


@pytest.mark.parametrize("cls", CLASSES, ids=lambda c: "".join(map(str, c)))
def test_small_class_against_oracle(gpu, port, cls):
    rng = np.random.default_rng(sum((i + 1) * l for i, l in enumerate(cls)) + 41)
    shells = random_shells(rng, list(cls), nprim=(1, 9), ncontr=(1, 3), box=1.5, log_exp=(-1.0, 1.5))
    b = make_basis(shells, port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    q = [(0, 1, 2, 3), (2, 3, 0, 1), (1, 0, 3, 2)]
    for B in ((0.1, -0.3, 0.25), (0.0, 0.0, 0.0)):
        f = _native.field6(B, (0.2, 0.1, -0.1))
        got = db.eri_batch(q, f)
        want, _ = port.eri_batch(b, q, port.field6(B, (0.2, 0.1, -0.1)), 16)
        err = float(np.max(np.abs(got - want)))
        print(f"{cls} B={B}: max |small kernel - oracle| = {err:.2e}, max |value| {np.abs(want).max():.2e}")
        assert err <= ATOL


def test_small_kernel_on_mixed_batch(gpu, port):
    """3000 mixed-class quartets: with and without the small-class kernel
    within 1e-11 of each other, and a sample against the oracle."""
    batch = workloads.mixed_batch(3000)
    sh = batch.shells
    coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
    db = _native.DeviceBasis(gpu, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)
    f = _native.field6(batch.B, batch.origin)
    try:
        gpu.set_small_max_l(-1)
        off = db.eri_batch(batch.quartets, f)
        gpu.set_small_max_l(2)
        on = db.eri_batch(batch.quartets, f)
    finally:
        gpu.set_small_max_l(2)
    diff = float(np.max(np.abs(on - off)))
    small = int(np.sum(sh.L[batch.quartets].sum(axis=1) <= 2))
    print(f"{small} of 3000 quartets on the small-class kernel; max |on - off| = {diff:.2e}")
    assert small > 500
    assert diff <= ATOL
    with pytest.raises(ValueError):
        gpu.set_small_max_l(3)
