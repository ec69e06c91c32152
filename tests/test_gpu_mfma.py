"""The E.R.E contraction on the FP64 matrix cores (gh_mfma.h) against the
oracle: every (LAB, LCD) instance in the orientation it is built for and
turned ((cd|ab) images of classes whose ket side is the larger), general
contractions, a field and zero field, at the 1e-11 bar of the FP64 path.
Reference contraction: src/eri.cpp:341-368."""

import numpy as np
import pytest

from giaoint_amd import _native
from oracle.pyoracle import make_basis
from tests.conftest import random_shells

pytestmark = pytest.mark.gpu

ATOL = 1e-11
# one class per instance (la, lb, lc, ld); the last ones put the larger pair on the ket side
CLASSES = [
    (1, 1, 2, 0), (2, 1, 1, 1), (3, 0, 1, 2), (2, 2, 0, 2), (3, 1, 2, 1), (2, 2, 3, 1),
    (3, 2, 1, 1), (2, 3, 3, 0), (3, 2, 2, 2), (3, 2, 2, 3), (3, 3, 2, 0), (3, 3, 1, 2),
    (3, 3, 2, 2), (4, 2, 3, 2), (3, 3, 3, 3), (4, 3, 3, 4), (4, 4, 4, 4),
    (2, 2, 3, 3), (1, 1, 2, 2), (0, 2, 4, 1),
]


@pytest.mark.parametrize("cls", CLASSES, ids=lambda c: "".join(map(str, c)))
def test_mfma_class_against_oracle(gpu, port, cls):
    rng = np.random.default_rng(sum((i + 1) * l for i, l in enumerate(cls)) + 7)
    big = sum(cls) >= 14
    # compact, diffuse shells: integrals of order 0.01..1, so the 1e-11 bar tests every element
    shells = random_shells(rng, list(cls), nprim=(1, 2) if big else (1, 3), ncontr=(1, 2) if big else (1, 3),
                           box=0.6, log_exp=(-1.0, 0.3))
    b = make_basis(shells, port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    q = [(0, 1, 2, 3), (2, 3, 0, 1)] if not big else [(0, 1, 2, 3)]
    gpu.set_mfma("all")
    try:
        for B in ((0.1, -0.3, 0.25), (0.0, 0.0, 0.0)):
            f = _native.field6(B, (0.2, 0.1, -0.1))
            got = db.eri_batch(q, f)
            want, _ = port.eri_batch(b, q, port.field6(B, (0.2, 0.1, -0.1)), 16)
            err = float(np.max(np.abs(got - want)))
            print(f"{cls} B={B}: max |MFMA - oracle| = {err:.2e}, max |value| {np.abs(want).max():.2e}")
            assert np.abs(want).max() > 1e-3
            assert err <= ATOL
    finally:
        gpu.set_mfma("default")


# The split-FP32 form (hi/lo FP32 products on v_mfma_f32_16x16x4f32, FP32
# sums per step of four compensated into FP64): instances (8, 8) and (6, 4),
# the latter also turned. It is a precision study, held to its own bound
# relative to the block's largest integral, not to the FP64 bar.
SPLIT_CLASSES = [(4, 4, 4, 4), (3, 3, 2, 2), (2, 2, 3, 3)]
SPLIT_REL = 1e-5


@pytest.mark.parametrize("cls", SPLIT_CLASSES, ids=lambda c: "".join(map(str, c)))
def test_split_fp32_class_against_oracle(gpu, port, cls):
    rng = np.random.default_rng(sum((i + 1) * l for i, l in enumerate(cls)) + 11)
    shells = random_shells(rng, list(cls), nprim=(1, 2), ncontr=(1, 2), box=0.6, log_exp=(-1.0, 0.3))
    b = make_basis(shells, port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    q = [(0, 1, 2, 3)]
    try:
        for B in ((0.1, -0.3, 0.25), (0.0, 0.0, 0.0)):
            f = _native.field6(B, (0.2, 0.1, -0.1))
            gpu.set_mfma("split32")
            got = db.eri_batch(q, f)
            gpu.set_mfma("all")
            fp64 = db.eri_batch(q, f)
            want, _ = port.eri_batch(b, q, port.field6(B, (0.2, 0.1, -0.1)), 16)
            err = float(np.max(np.abs(got - want)))
            scale = float(np.abs(want).max())
            print(f"{cls} B={B}: max |split - oracle| = {err:.2e} (FP64 MFMA {np.max(np.abs(fp64 - want)):.2e}), "
                  f"max |value| {scale:.2e}")
            assert scale > 1e-3
            assert np.max(np.abs(fp64 - want)) <= ATOL
            assert err <= SPLIT_REL * scale
            assert not np.array_equal(got, fp64)  # the split form really ran
    finally:
        gpu.set_mfma("default")


@pytest.mark.parametrize("cls", [(4, 4, 4, 4), (3, 3, 2, 2), (2, 1, 2, 1)], ids=lambda c: "".join(map(str, c)))
def test_mfma_compile_time_variants_bitwise(gpu, port, cls):
    """The four compile-time variants of the matrix-core kernels (GEMM step
    unroll, padded LDS rows) compute the same sums: bitwise identical
    integrals, and the oracle's at 1e-11."""
    rng = np.random.default_rng(sum((i + 1) * l for i, l in enumerate(cls)) + 23)
    shells = random_shells(rng, list(cls), nprim=(1, 2), ncontr=(1, 2), box=0.6, log_exp=(-1.0, 0.3))
    b = make_basis(shells, port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    q = [(0, 1, 2, 3), (2, 3, 0, 1)]
    f = _native.field6((0.1, -0.3, 0.25), (0.2, 0.1, -0.1))
    want, _ = port.eri_batch(b, q, port.field6((0.1, -0.3, 0.25), (0.2, 0.1, -0.1)), 16)
    gpu.set_mfma("all")
    try:
        got = {}
        for var in range(4):
            gpu.set_mfma_variant(var)
            got[var] = db.eri_batch(q, f)
        for var in range(1, 4):
            assert np.array_equal(got[var], got[0]), var
        err = float(np.max(np.abs(got[0] - want)))
        print(f"{cls}: four variants bitwise equal, max |MFMA - oracle| = {err:.2e}")
        assert err <= ATOL
        with pytest.raises(ValueError):
            gpu.set_mfma_variant(4)
    finally:
        gpu.set_mfma_variant(0)
        gpu.set_mfma("default")
