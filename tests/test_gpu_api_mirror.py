"""The reference-named Python layer on the GPU against the oracle: same
results as the C restatement for eri_shell, _eri_batch (layout, screening
counts), _eri_full (with and without screening), boys, primitive_eri, the
selected Cholesky integrals and cholesky_eri."""

import numpy as np
import pytest

from giaoint_amd import _native, cholesky, core, eri, workloads
from oracle.pyoracle import make_basis
from tests.conftest import random_shells

pytestmark = pytest.mark.gpu

ATOL = 1e-11
B, O = (0.1, -0.3, 0.25), (0.2, 0.1, -0.1)


def mirror_basis(shells):
    return core.Basis([core.Shell(c, l, e, k) for c, l, e, k in shells])


def test_blocks_batch_and_full_tensor(port):
    rng = np.random.default_rng(21)
    shells = random_shells(rng, [0, 1, 2, 1])
    mb = mirror_basis(shells)
    ob = make_basis(shells, port)
    field = core.MagneticField(B, O)
    f6 = port.field6(B, O)
    blk = core.eri_shell(*mb.shells[:2], *mb.shells[2:], field=field)
    want, _ = port.eri_batch(ob, [(0, 1, 2, 3)], f6, 4)
    assert np.max(np.abs(blk.reshape(-1) - want)) <= ATOL
    out = np.zeros_like(blk)
    assert core.eri_shell(*mb.shells, field=field, out=out) is out
    with pytest.raises(ValueError, match="incorrect shape"):
        core.eri_shell(*mb.shells, field=field, out=np.zeros((1, 1, 1, 1), dtype=np.complex128))
    qs = [(0, 1, 2, 3), (3, 2, 1, 0), (1, 1, 2, 2)]
    q, shapes, offsets, values, requested, screened = core._eri_batch(mb, qs, field=field)
    assert (requested, screened) == (3, 0) and len(offsets) == 4 and offsets[-1] == len(values)
    want, woff = port.eri_batch(ob, qs, f6, 4)
    np.testing.assert_array_equal(offsets, woff.astype(np.int64))
    assert np.max(np.abs(values - want)) <= ATOL
    full = core._eri_full(mb, field=field)
    assert np.max(np.abs(full - port.eri_tensor(ob, f6, 4))) <= ATOL
    # screening: blocks below the bound product are left out / zero
    bounds = core._EriSchwarzBounds(mb, field=field)
    Q = bounds.values
    np.testing.assert_allclose(Q, port.schwarz_bounds(ob, f6), rtol=1e-12)
    thr = float(np.median(np.outer(Q, Q)))
    batches = list(eri.eri_batches(mb, field=field, screening_threshold=thr))
    canon = list(eri._quartets(len(shells), None))
    assert sum(b.requested_count for b in batches) == len(canon)
    kept = [tuple(x) for b in batches for x in b.quartets.tolist()]
    assert kept == [c for c in canon if Q[c[0], c[1]] * Q[c[2], c[3]] >= thr]
    assert sum(b.screened_count for b in batches) == len(canon) - len(kept)
    b0 = batches[0]
    assert b0.block(0).shape == tuple(b0.shapes[0])
    scr = eri.eri(mb, field=field, storage="full", max_bytes=full.nbytes, screening_threshold=thr)
    assert np.max(np.abs(scr - full)) <= 2 * thr  # each dropped block is bounded by its Schwarz product
    with pytest.raises(ValueError, match="do not match"):
        core._eri_batch(mb, qs, field=core.MagneticField(), screening_threshold=thr, bounds=bounds)


def test_boys_and_primitive(port):
    v, d = core.boys_with_diagnostics(20 - 35j, 8)
    ref, info = port.boys(20 - 35j, 8)
    assert d["region"] == core.BoysRegion(info["region"]) and d["scaled"] is False
    assert np.all(np.abs(v - ref) <= 3e-14 + 8e-13 * np.abs(ref))
    np.testing.assert_array_equal(core.boys(0.0, 3), [1.0, 1 / 3, 1 / 5, 1 / 7])
    with pytest.raises(core.BoysNumericalError):
        core.boys(-400 + 10j, 4)
    p = [core.PrimitiveGaussian(0.8 + 0.1 * i, (0.1 * i, -0.2, 0.3 * i), angular=(i % 2, 0, 1)) for i in range(4)]
    got = core.primitive_eri(*p, field=core.MagneticField(B, O))
    want = port.primitive_eri([{"exponent": x.exponent, "center": x.center, "angular": x.angular} for x in p],
                              port.field6(B, O))
    assert abs(got - want) <= ATOL


def test_selected_integrals_and_cholesky_eri(port):
    sh, Bf, Of = workloads.cholesky_atoms(natoms=2)
    shells = sh.shell_list()
    mb = mirror_basis(shells)
    ob = make_basis(shells, port)
    field = core.MagneticField(Bf, Of)
    f6 = port.field6(Bf, Of)
    pairs = [(a, b) for a in range(len(shells)) for b in range(a, len(shells))]
    assert np.max(np.abs(core._cholesky_diagonal(mb, pairs, field=field) - port.cholesky_diagonal(ob, pairs, f6, 4))) \
        <= ATOL
    piv = [(0, 3), (2, 5)]
    assert np.max(np.abs(core._cholesky_columns(mb, piv[:1], field=field)
                         - port.cholesky_columns(ob, piv[:1], f6, 4))) <= ATOL
    real = core._cholesky_columns(mb, piv[:1], field=None)
    assert real.dtype == np.float64
    res = cholesky.cholesky_eri(mb, field=field, threshold=1e-8)
    assert res.max_residual_diagonal <= 1e-8 and res.threshold == 1e-8
    db = _native.DeviceBasis(_native.Context(0), ob.centers, ob.L, ob.nprim, ob.ncontr, ob.exps, ob.coefs)
    run = _native.cholesky_run(db, f6, 1e-8)
    np.testing.assert_array_equal(res.pivots, run.pivots())
    np.testing.assert_array_equal(res.vectors, run.vectors())
    full = port.eri_tensor(ob, f6, 4)
    assert np.max(np.abs(res.reconstruct() - full)) <= 1e-8
    row = mb.ao_count**2 * 16 + 16
    with pytest.raises(MemoryError, match=f"exceeding max_bytes={3 * row}"):
        cholesky.cholesky_eri(mb, field=field, max_bytes=3 * row)
    with pytest.raises(MemoryError, match="exceeding max_bytes=0"):
        cholesky.cholesky_eri(mb, field=field, max_bytes=0)


def test_molecule_entry_points(gpu, port):
    """Molecule (molecule.py:285-450) on water cc-pVDZ from the committed BSE
    JSON: the full tensor against the checker's, and the Cholesky factors
    (general contractions, complex field) reconstructing it within the
    threshold."""
    from giaoint_amd.molecule import Molecule
    from oracle.pyoracle import make_basis

    atoms = [("O", (0.0, 0.0, 0.1173)), ("H", (0.0, 0.7572, -0.4692)), ("H", (0.0, -0.7572, -0.4692))]
    mol = Molecule(atoms, workloads.DATA / "cc-pvdz-h-o.json", field=core.MagneticField((0.0, 0.0, 0.1)))
    T = mol.eri()
    sh, B, O = workloads.water_ccpvdz()
    b = make_basis(sh.shell_list(), port)
    want = port.eri_tensor(b, port.field6(B, O), threads=8)
    err = float(np.max(np.abs(T - want)))
    print(f"Molecule.eri max |err| {err:.2e}")
    assert err <= 1e-11
    ch = mol.cholesky_eri(threshold=1e-6)
    recon = ch.reconstruct()
    gap = float(np.max(np.abs(recon - want)))
    print(f"Molecule.cholesky_eri rank {len(ch.pivots)}, max |ERI - LL*| {gap:.2e}")
    assert gap <= 1e-6 + 1e-11
