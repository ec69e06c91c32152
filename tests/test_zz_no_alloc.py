"""giaohip_eri_batch_noalloc: the ERI batch into caller-provided device
memory without any device allocation (its work memory is the library's
static per-device array), against giaohip_eri_batch into device memory and
against the C oracle, on two calls in a row, the second smaller."""

import ctypes as C
import itertools

import numpy as np
import pytest

from giaoint_amd import _native
from oracle.pyoracle import make_basis
from tests.conftest import random_shells

pytestmark = pytest.mark.gpu

ATOL = 1e-11


def _noalloc(ctx, db, quartets, field, out_ptr):
    q = np.ascontiguousarray(np.asarray(quartets, dtype=np.uint32).reshape(-1, 4))
    f = np.ascontiguousarray(field, dtype=np.float64)
    ctx.check(ctx.lib.giaohip_eri_batch_noalloc(ctx.handle, db.handle, q.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                C.c_size_t(len(q)), f.ctypes.data_as(C.POINTER(C.c_double)), None,
                                                C.c_void_p(out_ptr)))


def _device_values(ctx, ptr, count):
    out = np.empty(count, dtype=np.complex128)
    return ctx.to_host(ptr, out)


def test_noalloc_matches_eri_batch(gpu, port):
    rng = np.random.default_rng(11)
    shells = random_shells(rng, [0, 1, 2, 3, 2, 1, 0, 2], nprim=(1, 4), ncontr=(1, 3))
    b = make_basis(shells, port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    field = _native.field6((0.1, -0.2, 0.3), (0.2, 0.0, -0.1))
    big = np.array(list(itertools.product(range(8), repeat=4)), dtype=np.uint32)
    small = big[rng.permutation(len(big))[:300]]
    sizes_big = db.quartet_sizes(big)
    sizes_small = db.quartet_sizes(small)
    total = int(sizes_big.sum())
    ref_ptr = gpu.alloc(total * 16)
    out_ptr = gpu.alloc(total * 16)
    try:
        lib = gpu.lib
        before = lib.giaohip_device_allocations()
        print(f"noalloc call 1: {len(big)} quartets, {total} values", flush=True)
        _noalloc(gpu, db, big, field, out_ptr)
        got_big = _device_values(gpu, out_ptr, total)
        print(f"noalloc call 2: {len(small)} quartets", flush=True)
        _noalloc(gpu, db, small, field, out_ptr)
        got_small = _device_values(gpu, out_ptr, int(sizes_small.sum()))
        assert lib.giaohip_device_allocations() == before
        db.eri_batch(big, field, device_out=ref_ptr)
        want_big = _device_values(gpu, ref_ptr, total)
        db.eri_batch(small, field, device_out=ref_ptr)
        want_small = _device_values(gpu, ref_ptr, int(sizes_small.sum()))
        assert float(np.max(np.abs(got_big - want_big))) <= ATOL
        assert float(np.max(np.abs(got_small - want_small))) <= ATOL
        oracle_small, _ = port.eri_batch(b, small, field, 8)
        assert float(np.max(np.abs(got_small - oracle_small))) <= ATOL
    finally:
        gpu.free(ref_ptr)
        gpu.free(out_ptr)


def test_noalloc_rejects_oversized_quartets(gpu, port):
    """A shell quartet whose seeds exceed the static scratch is an error, not
    an allocation: four s shells of 30 primitives (810,000 primitive quartets)."""
    rng = np.random.default_rng(12)
    shells = [(rng.uniform(-1, 1, 3), 0, 10 ** rng.uniform(-1, 2, 30), rng.normal(size=(1, 30))) for _ in range(4)]
    b = make_basis(shells, port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    ptr = gpu.alloc(16)
    try:
        with pytest.raises(ValueError, match="exceeds the allocation-free ERI scratch"):
            _noalloc(gpu, db, [(0, 1, 2, 3)], _native.field6((0, 0, 0.1), (0, 0, 0)), ptr)
    finally:
        gpu.free(ptr)
