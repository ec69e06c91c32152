"""Launch forms and autotuned launch parameters give the same integrals.

* The generic transform's three launch forms -- one workgroup per item and
  LDS-size bucket, a persistent device work queue per bucket, one persistent
  launch over every item -- give bitwise identical integrals on the
  mixed-class workload (l <= 3, degree 1..8), which match the oracle on a
  sample at 1e-11.
* Every workgroup size of the class-specialized (dd|dd) kernel (the bra-pair
  tile) matches the oracle at 1e-11; chunk sizes of the batched path give
  bitwise identical integrals.
* The autotuner, run on a small ERI batch and on a short Cholesky run,
  returns only accepted values; Cholesky pivots are bitwise equal across
  the block sizes it tries.

Each test uses a context of its own, so that options do not leak into other
tests.
"""

import numpy as np
import pytest

from giaoint_amd import _native, autotune, workloads

pytestmark = pytest.mark.gpu

ATOL = 1e-11


def _device(ctx, batch):
    sh = batch.shells
    coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
    return _native.DeviceBasis(ctx, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)


def _oracle_rows(port, batch, rows):
    from oracle.pyoracle import make_basis

    sh = batch.shells
    sub = sh.subset(np.concatenate([batch.quartets[i] for i in rows]).astype(np.int64))
    b = make_basis(sub.shell_list(), port)
    q = np.arange(4 * len(rows), dtype=np.uint32).reshape(-1, 4)
    vals, _ = port.eri_batch(b, q, port.field6(batch.B, batch.origin), 16)
    return vals


def _rows_of(db, batch, values, rows):
    sizes = db.quartet_sizes(batch.quartets)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    return np.concatenate([values[offs[i]:offs[i] + sizes[i]] for i in rows])


@pytest.mark.parametrize("variant", ["plain", "batched"])
def test_queue_forms_bitwise(port, variant):
    ctx = _native.Context(0)
    batch = workloads.mixed_batch(3000)
    db = _device(ctx, batch)
    f = _native.field6(batch.B, batch.origin)
    ctx.set_variant(variant)
    got = {}
    for form in ("none", "buckets", "single", "buckets"):
        print(f"queue form {form}, variant {variant}, 3000 mixed quartets", flush=True)
        ctx.set_eri_queue(form)
        got[form] = db.eri_batch(batch.quartets, f)
    for form in ("buckets", "single"):
        assert np.array_equal(got[form], got["none"]), form
    rows = np.arange(0, 3000, 15)
    err = np.abs(_rows_of(db, batch, got["buckets"], rows) - _oracle_rows(port, batch, rows))
    print(f"work-queue launch vs oracle on {len(rows)} quartets: max |err| {err.max():.2e}")
    assert err.max() <= ATOL


def test_wave_workgroups_bitwise(port):
    """One-wavefront workgroups for the generic transform's small items give
    the integrals of four-wavefront workgroups bitwise (the per-output sums
    are the same; only the ket components per item and the occupancy
    change)."""
    ctx = _native.Context(0)
    batch = workloads.mixed_batch(3000)
    db = _device(ctx, batch)
    f = _native.field6(batch.B, batch.origin)
    ctx.set_variant("plain")
    got = {}
    for lmax in (-1, 7, 12, 3):
        print(f"one-wavefront items up to L = {lmax}, 3000 mixed quartets", flush=True)
        ctx.set_wave_max_l(lmax)
        got[lmax] = db.eri_batch(batch.quartets, f)
    for lmax in (7, 12, 3):
        assert np.array_equal(got[lmax], got[-1]), lmax
    rows = np.arange(0, 3000, 15)
    err = np.abs(_rows_of(db, batch, got[12], rows) - _oracle_rows(port, batch, rows))
    print(f"one-wavefront items vs oracle on {len(rows)} quartets: max |err| {err.max():.2e}")
    assert err.max() <= ATOL


def test_fast_waves_tiles(port):
    ctx = _native.Context(0)
    batch = workloads.dddd_batch(96)
    db = _device(ctx, batch)
    f = _native.field6(batch.B, batch.origin)
    rows = np.arange(0, 96, 3)
    want = _oracle_rows(port, batch, rows)
    for waves in autotune.FAST_WAVES:
        print(f"(dd|dd) class kernel, {waves or 8} waves per workgroup", flush=True)
        ctx.set_fast_waves(waves)
        err = np.abs(_rows_of(db, batch, db.eri_batch(batch.quartets, f), rows) - want)
        print(f"  max |err| {err.max():.2e}")
        assert err.max() <= ATOL


def test_chunk_sizes_bitwise():
    """The batched path's chunk size only regroups independent quartets."""
    ctx = _native.Context(0)
    batch = workloads.dddd_batch(500)
    db = _device(ctx, batch)
    f = _native.field6(batch.B, batch.origin)
    got = {}
    for mib in (16, 64, 2048):
        print(f"(dd|dd) 500 quartets, chunk {mib} MiB", flush=True)
        ctx.set_chunk_mib(mib)
        got[mib] = db.eri_batch(batch.quartets, f)
        print(f"  launches {ctx.stats()['contract_launches']}")
    assert np.array_equal(got[16], got[2048]) and np.array_equal(got[64], got[2048])


def test_pipeline_forms_bitwise():
    """Producers beside the previous chunk's contraction or serially on the
    context stream: the same kernels on the same data, bitwise equal, over
    several chunks of (dd|dd) and a sample of the mixed classes."""
    ctx = _native.Context(0)
    ctx.set_chunk_mib(16)
    for batch in (workloads.dddd_batch(600), workloads.mixed_batch(300)):
        db = _device(ctx, batch)
        f = _native.field6(batch.B, batch.origin)
        got = {}
        for mode in ("overlap", "serial"):
            ctx.set_pipeline(mode)
            got[mode] = db.eri_batch(batch.quartets, f)
            print(f"{batch.description}: {mode}, {ctx.stats()['contract_launches']} contraction launches", flush=True)
        assert np.array_equal(got["overlap"], got["serial"])
    with pytest.raises(ValueError):
        ctx.check(ctx.lib.giaohip_set_option(ctx.handle, 8, 7))


def test_fast_waves_rejects_out_of_range():
    ctx = _native.Context(0)
    for bad in (1, 3, 9):
        with pytest.raises(ValueError):
            ctx.set_fast_waves(bad)


def test_autotune_eri_small_batch():
    ctx = _native.Context(0)
    batch = workloads.mixed_batch(2000)
    db = _device(ctx, batch)
    f = _native.field6(batch.B, batch.origin)
    choice = autotune.tune_eri(db, batch.quartets, f, reps=1, check_quartets=500)
    print(choice.as_dict())
    assert choice.values["fast_waves"] in autotune.FAST_WAVES
    assert choice.values["eri_queue"] in autotune.ERI_QUEUE
    queue_trials = [t for t in choice.trials if t.knob == "eri_queue"]
    assert len(queue_trials) == len(autotune.ERI_QUEUE)
    # the launch forms do not change a single bit
    assert all(t.accepted and t.max_abs_diff == 0.0 for t in queue_trials)
    assert all(t.max_abs_diff <= ATOL for t in choice.trials if t.knob == "fast_waves")


def test_autotune_cholesky_blocks_bitwise():
    ctx = _native.Context(0)
    sh, B, O = workloads.cholesky_atoms(natoms=3)
    coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
    db = _native.DeviceBasis(ctx, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)
    choice = autotune.tune_cholesky(db, _native.field6(B, O), 1e-8, 200)
    print(choice.as_dict())
    assert all(t.accepted for t in choice.trials)
    assert choice.values["cholesky_block"] in autotune.CHOLESKY_BLOCK
