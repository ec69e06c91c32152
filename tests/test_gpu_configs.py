"""Samples of the mixed-class and high-l workloads against the oracle.

* Mixed classes l <= 3 with contraction degree U{1..8} and class weights
  (1 + sum l)^-2 (the 2M-quartet workload's generator at 3000 quartets):
  the batched variant (class-specialized kernels where they exist, the plain
  kernel elsewhere) agrees with the plain variant on every quartet, and with
  the oracle on 300 of them.
* (gg|gg) and (ff|dd) with exponents log-uniform in [0.05, 5e3] and degree
  2: max and RMS error against the oracle (the FP64 path; bar 1e-11).
"""

import numpy as np
import pytest

from giaoint_amd import _native, workloads

pytestmark = pytest.mark.gpu

ATOL = 1e-11


def _device(ctx, batch):
    sh = batch.shells
    coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
    return _native.DeviceBasis(ctx, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)


def _oracle(port, batch, rows):
    from oracle.pyoracle import make_basis

    sh = batch.shells
    sub = sh.subset(np.concatenate([batch.quartets[i] for i in rows]).astype(np.int64))
    b = make_basis(sub.shell_list(), port)
    q = np.arange(4 * len(rows), dtype=np.uint32).reshape(-1, 4)
    vals, _ = port.eri_batch(b, q, port.field6(batch.B, batch.origin), 16)
    return vals


def _blocks(db, batch):
    sizes = db.quartet_sizes(batch.quartets)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    return offs, sizes


def test_mixed_class_sample(gpu, port):
    batch = workloads.mixed_batch(3000)
    db = _device(gpu, batch)
    f = _native.field6(batch.B, batch.origin)
    try:
        gpu.set_variant("plain")
        plain = db.eri_batch(batch.quartets, f)
    finally:
        gpu.set_variant("batched")
    fast = db.eri_batch(batch.quartets, f)
    assert np.max(np.abs(fast - plain)) <= ATOL
    rows = np.arange(0, 3000, 10)
    offs, sizes = _blocks(db, batch)
    got = np.concatenate([fast[offs[i]:offs[i] + sizes[i]] for i in rows])
    want = _oracle(port, batch, rows)
    err = np.abs(got - want)
    print(f"mixed sample: {len(rows)} quartets, max |err| {err.max():.2e}")
    assert err.max() <= ATOL


def test_high_l_precision_samples(gpu, port):
    g, fd = workloads.precision_batches(n_gggg=4, n_ffdd=100)
    for batch in (g, fd):
        db = _device(gpu, batch)
        f = _native.field6(batch.B, batch.origin)
        got = db.eri_batch(batch.quartets, f)
        want = _oracle(port, batch, np.arange(len(batch.quartets)))
        err = np.abs(got - want)
        rms = float(np.sqrt(np.mean(err**2)))
        print(f"{batch.description}: max |err| {err.max():.2e}, rms {rms:.2e}, max |value| {np.abs(want).max():.2e}")
        assert err.max() <= ATOL


def test_high_l_mfma_variant_errors(gpu, port):
    """The FP64 matrix-core contraction on the precision workloads: within
    1e-10 of the FP64 VALU path on 200 (gg|gg) and 2000 (ff|dd) quartets,
    and of the oracle on 2 and 50 of them; max and RMS errors printed."""
    g, fd = workloads.precision_batches(n_gggg=200, n_ffdd=2000)
    for batch, n_oracle in ((g, 2), (fd, 50)):
        db = _device(gpu, batch)
        f = _native.field6(batch.B, batch.origin)
        try:
            gpu.set_mfma("off")
            valu = db.eri_batch(batch.quartets, f)
            gpu.set_mfma("all")
            mfma = db.eri_batch(batch.quartets, f)
        finally:
            gpu.set_mfma("default")
        err = np.abs(mfma - valu)
        offs, sizes = _blocks(db, batch)
        want = _oracle(port, batch, np.arange(n_oracle))
        oerr = np.abs(mfma[: offs[n_oracle]] - want)
        print(f"{batch.description}: MFMA vs FP64 VALU max {err.max():.2e} rms {np.sqrt(np.mean(err**2)):.2e}; "
              f"vs oracle ({n_oracle} quartets) max {oerr.max():.2e}; max |value| {np.abs(valu).max():.2e}")
        assert err.max() <= 1e-10
        assert oerr.max() <= 1e-10
