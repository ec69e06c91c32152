"""GPU ERI path (through the C ABI) against the oracle.

Bar: |GPU - oracle| <= 1e-11 absolute per complex ERI element (FP64 path);
Boys values within the reference's own Boys test tolerance.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from giaoint_amd import _native, workloads
from oracle.pyoracle import make_basis
from tests.conftest import random_shells

pytestmark = pytest.mark.gpu

ATOL = 1e-11
FIELD = ((0.1, -0.3, 0.25), (0.2, 0.1, -0.1))
RECORDS = [json.loads(x) for x in (Path(__file__).parent / "golden" / "eri_recursive_m8.jsonl").read_text().splitlines()]
BOYS_ARGS = [0.0, 1e-14 + 1e-14j, 0.749999, 0.750001, 1 + 4j, 20 - 35j, -1 + 2j, -30 - 12j, -100 + 20j, 123.999,
             124.001, 130 + 10j, 159.999j, 1000 - 25j]


def device_basis(ctx, b):
    return _native.DeviceBasis(ctx, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)


@pytest.mark.parametrize("nmax", [0, 3, 12, 32])
def test_boys_against_oracle(gpu, port, nmax):
    vals, region, _, _ = gpu.boys(BOYS_ARGS, nmax)
    for i, z in enumerate(BOYS_ARGS):
        ref, info = port.boys(z, nmax)
        assert region[i] == info["region"]
        assert np.all(np.abs(vals[i] - ref) <= 3e-14 + 8e-13 * np.abs(ref))


def test_boys_scaled_and_errors(gpu, port):
    for z, scaled in ((-0.3 + 0.2j, True), (-20 + 5j, True), (-120 - 8j, False), (40 + 9j, True)):
        v, _, _, _ = gpu.boys([z], 16, scaled=scaled)
        r, _ = port.boys(z, 16, scaled)
        assert np.all(np.abs(v[0] - r) <= 2e-14 + 8e-13 * np.abs(r))
    # the reference's adaptive quadrature gives up on exp(z(1-t^2)) at z = 40+90i
    from oracle.pyoracle import CheckerError

    with pytest.raises(CheckerError, match="did not reach tolerance"):
        port.boys(40 + 90j, 16, True)
    with pytest.raises(_native.BoysNumericalError, match="did not reach tolerance"):
        gpu.boys([40 + 90j], 16, scaled=True)
    with pytest.raises(_native.BoysNumericalError, match="outside the verified"):
        gpu.boys([10 + 500j], 4)
    with pytest.raises(ValueError, match="through 32"):
        gpu.boys([1.0], 33)


def test_primitive_eri_frozen_vectors(gpu):
    for r in RECORDS:
        for field, expected in zip([(0, 0, 0), r["field"]], r["expected"]):
            v = gpu.primitive_eri(r["primitives"], _native.field6(field, r["gauge"]))
            assert abs(v - complex(*expected)) <= 5e-14 + 5e-13 * abs(complex(*expected))


@pytest.mark.parametrize("variant", ["batched", "plain"])
@pytest.mark.parametrize("field", [FIELD, ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0))], ids=["finite", "zero"])
def test_blocks_all_classes_to_g(gpu, port, field, variant):
    """Assorted classes; the batched variant reaches its specialized kernels
    directly ((dd|dd), (pp|pp), (ss|ss)), through pair exchange ((ff|dd) ->
    (dd|ff)) and through conjugate reversal ((sp|sp) -> (ps|ps)*)."""
    rng = np.random.default_rng(11)
    ls = [0, 1, 2, 3, 4, 0, 1, 2, 3, 4, 2, 2]
    b = make_basis(random_shells(rng, ls, nprim=(1, 3), ncontr=(1, 3)), port)
    db = device_basis(gpu, b)
    qs = [(0, 0, 0, 0), (1, 0, 2, 3), (2, 2, 2, 2), (3, 2, 1, 0), (4, 3, 2, 1), (3, 3, 2, 2), (4, 4, 0, 0),
          (0, 4, 1, 9), (5, 6, 7, 8), (9, 8, 7, 6), (2, 7, 2, 7), (1, 1, 1, 1), (3, 3, 3, 3), (0, 1, 5, 6),
          (1, 0, 6, 5), (3, 8, 10, 11), (8, 3, 2, 7), (2, 11, 7, 10), (10, 11, 3, 8)]
    f = _native.field6(*field)
    gpu.set_variant(variant)
    try:
        got = db.eri_batch(qs, f)
    finally:
        gpu.set_variant("batched")
    ref, _ = port.eri_batch(b, qs, f, threads=8)
    assert np.max(np.abs(got - ref)) <= ATOL


def test_gggg_block(gpu, port):
    rng = np.random.default_rng(12)
    b = make_basis(random_shells(rng, [4, 4, 4, 4], nprim=(2, 3), ncontr=(1, 2)), port)
    db = device_basis(gpu, b)
    f = _native.field6(*FIELD)
    got = db.eri_batch([(0, 1, 2, 3)], f)
    ref, _ = port.eri_batch(b, [(0, 1, 2, 3)], f, threads=1)
    assert np.max(np.abs(got - ref)) <= ATOL


def test_packed_layout_and_errors(gpu, port):
    rng = np.random.default_rng(13)
    b = make_basis(random_shells(rng, [0, 1, 2]), port)
    db = device_basis(gpu, b)
    f = _native.field6(*FIELD)
    single = [db.eri_batch([q], f) for q in [(0, 1, 2, 0), (2, 2, 1, 1)]]
    both = db.eri_batch([(0, 1, 2, 0), (2, 2, 1, 1)], f)
    np.testing.assert_array_equal(both, np.concatenate(single))
    with pytest.raises(IndexError, match="out of range"):
        db.eri_batch([(0, 1, 2, 3)], f)
    assert db.eri_batch(np.zeros((0, 4)), f).size == 0


def test_combined_order_limit(gpu, port):
    shells = [(np.zeros(3), 9, np.array([1.0]), np.array([[1.0]]))] * 3 + [(np.ones(3), 8, np.array([1.0]), np.array([[1.0]]))]
    b = make_basis(shells, port)
    db = device_basis(gpu, b)
    with pytest.raises(ValueError, match="through 32"):
        db.eri_batch([(0, 1, 2, 3)], _native.field6(*FIELD))


def test_water_ccpvdz_tensor(gpu, port):
    sh, B, O = workloads.water_ccpvdz()
    b = make_basis(sh.shell_list(), port)
    db = device_basis(gpu, b)
    f = _native.field6(B, O)
    got = db.eri_tensor(f)
    ref = port.eri_tensor(b, f, threads=16)
    assert got.shape == (25, 25, 25, 25)
    assert np.max(np.abs(got - ref)) <= ATOL


def test_dddd_batch_sample_and_symmetries():
    """Full 200k (dd|dd) workload: pair exchange and conjugate reversal hold
    for every block; a seeded sample matches the oracle."""
    from oracle import pyoracle

    batch = workloads.dddd_batch()
    sh = batch.shells
    coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
    ctx = _native.Context(0)
    db = _native.DeviceBasis(ctx, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)
    f = _native.field6(batch.B, batch.origin)
    q = batch.quartets
    a = db.eri_batch(q, f).reshape(len(q), 6, 6, 6, 6)
    swapped = db.eri_batch(q[:, [2, 3, 0, 1]], f).reshape(len(q), 6, 6, 6, 6)
    assert np.max(np.abs(a - swapped.transpose(0, 3, 4, 1, 2))) <= ATOL
    del swapped
    rev = db.eri_batch(q[:, [1, 0, 3, 2]], f).reshape(len(q), 6, 6, 6, 6)
    assert np.max(np.abs(a - rev.transpose(0, 2, 1, 4, 3).conj())) <= ATOL
    del rev
    port = pyoracle.load("port")
    pick = np.random.default_rng(99).choice(len(q), 24, replace=False)
    sub = batch.shells.subset(q[pick].reshape(-1))
    b = make_basis(sub.shell_list(), port)
    ref, _ = port.eri_batch(b, np.arange(4 * len(pick)).reshape(-1, 4), f, threads=16)
    assert np.max(np.abs(a[pick].reshape(-1) - ref)) <= ATOL


def test_schwarz_bounds_and_screening(gpu, port):
    """Device Schwarz factors against the oracle's (eri.cpp:587-620), and the
    screening step of evaluate_eri_shell_quartets (eri.cpp:652-688): kept
    quartets in input order, validation before screening."""
    rng = np.random.default_rng(11)
    shells = random_shells(rng, [0, 1, 2, 3, 1, 0, 2])
    b = make_basis(shells, port)
    db = device_basis(gpu, b)
    for f in (port.field6(*FIELD), port.field6()):
        Q = db.schwarz_bounds(f)
        Qp = port.schwarz_bounds(b, f)
        np.testing.assert_allclose(Q, Qp, rtol=1e-12, atol=0)
        assert np.array_equal(Q, Q.T)
    ns = len(shells)
    qs = np.array([(a, c, d, e) for a in range(ns) for c in range(ns) for d in range(ns) for e in range(ns)],
                  dtype=np.uint32)
    prod = Qp[qs[:, 0], qs[:, 1]] * Qp[qs[:, 2], qs[:, 3]]
    thr = float(np.median(prod))
    kept = db.screen(qs, Qp, thr)
    np.testing.assert_array_equal(kept, qs[prod >= thr])
    assert 0 < len(kept) < len(qs)
    np.testing.assert_array_equal(db.screen(qs, None, 0.0), qs)
    with pytest.raises(IndexError):
        db.screen(np.array([[0, 0, 0, ns]]), Qp, thr)
    with pytest.raises(ValueError, match="finite and non-negative"):
        db.screen(qs, Qp, -1.0)
    with pytest.raises(ValueError, match="requires Schwarz bounds"):
        db.screen(qs, None, 1e-8)
    # the kept quartets evaluate as usual
    f = port.field6(*FIELD)
    got = db.eri_batch(kept[:40], f)
    want, _ = port.eri_batch(b, kept[:40], f, 4)
    assert np.max(np.abs(got - want)) <= ATOL


def test_host_output_groups_and_scattered_offsets(gpu):
    """Outputs written to host memory go through device staging groups that
    are copied out while the next group is computed. With a 1 MiB staging
    budget, 2000 (dd|dd) quartets (41 MB of output) take ~40 groups; the host
    result must equal the device-resident one bitwise, also when the caller
    asks for blocks at scattered offsets (giaohip_eri_batch's out_offsets)."""
    import ctypes as C

    batch = workloads.dddd_batch(2000, seed=7)
    sh = batch.shells
    coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
    db = _native.DeviceBasis(gpu, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)
    f = _native.field6(batch.B, batch.origin)
    sizes = db.quartet_sizes(batch.quartets)
    total = int(sizes.sum())
    dev = gpu.alloc(total * 16)
    try:
        db.eri_batch(batch.quartets, f, device_out=dev)
        want = gpu.to_host(dev, np.empty(total, dtype=np.complex128))
    finally:
        gpu.free(dev)
    gpu.set_staging_mib(1)
    try:
        got = db.eri_batch(batch.quartets, f)
        np.testing.assert_array_equal(got, want)
        # every block 7 elements apart, in reverse order
        q = np.ascontiguousarray(batch.quartets, dtype=np.uint32)
        gap = 7
        offs = np.zeros(len(q), dtype=np.uint64)
        pos = 0
        for i in reversed(range(len(q))):
            offs[i] = pos
            pos += int(sizes[i]) + gap
        out = np.full(pos, np.nan + 0j, dtype=np.complex128)
        gpu.check(gpu.lib.giaohip_eri_batch(
            gpu.handle, db.handle, q.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_size_t(len(q)),
            f.ctypes.data_as(C.POINTER(C.c_double)), offs.ctypes.data_as(C.POINTER(C.c_uint64)),
            out.ctypes.data_as(C.c_void_p), 0))
        starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        for i in range(len(q)):
            o, n = int(offs[i]), int(sizes[i])
            np.testing.assert_array_equal(out[o:o + n], want[starts[i]:starts[i] + n])
        assert np.isnan(out[int(offs[0]) + int(sizes[0]):int(offs[0]) + int(sizes[0]) + gap].real).all()
    finally:
        gpu.set_staging_mib(1024)
