"""ERI centre-derivative batches on the GPU (compute_eri_derivative_blocks,
derivatives.cpp:1389-1415; _eri_derivative_batch, bindings.cpp:2255-2304;
eri_derivative_batches, eri.py:220-288) against the oracle's restatement.

Bar: |GPU - oracle| <= 1e-11 x the largest |value| of the batch per complex
element (FP64; the device contracts the shifted-momentum integrals before
combining them, the reference combines per primitive quartet). Derivative
Schwarz bounds within 1e-12 relative. At the cc-pVTZ water size, the
translation sum rule: the four slot derivatives along an axis add up to
i/2 (B x (A-B) + B x (C-D)) (ab|cd). The streaming, screening and error
cases follow tests/python/test_eri_derivative_batches.py.
"""

from itertools import product

import numpy as np
import pytest

from giaoint_amd import _native, core, eri, workloads
from giaoint_amd.molecule import Molecule
from oracle.pyoracle import make_basis
from tests.conftest import random_shells

pytestmark = pytest.mark.gpu

FIELDS = [None, core.MagneticField((0.17, -0.11, 0.23), (0.2, -0.3, 0.1))]
WATER_BOHR = [("O", (0.0, 0.1, 0.2)), ("H", (0.1, 1.4, -0.9)), ("H", (-0.2, -1.3, -1.0))]


def _f6(field):
    return np.array(field._field6() if field is not None else (0.0,) * 6)


def _core_basis(raw):
    return core.Basis([core.Shell(c, l, e, k) for c, l, e, k in raw])


@pytest.mark.parametrize("field", FIELDS)
def test_blocks_match_oracle(gpu, port, rng, field):
    raw = random_shells(rng, [0, 1, 2, 3, 1, 2], nprim=(1, 3), ncontr=(1, 3))
    b = make_basis(raw, port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    qs = [(0, 1, 2, 3), (3, 3, 2, 2), (5, 4, 1, 0), (2, 2, 2, 2), (1, 0, 1, 0), (4, 3, 5, 1), (0, 0, 0, 0),
          (3, 2, 3, 2)]
    got = db.eri_derivative_batch(qs, _f6(field))
    ref, offs = port.eri_derivative_blocks(b, qs, _f6(field), threads=8)
    scale = float(np.max(np.abs(ref)))
    err = float(np.max(np.abs(got - ref)))
    print(f"derivative blocks, field {field is not None}: max |GPU - oracle| = {err:.3e}, scale {scale:.3e}",
          flush=True)
    assert err <= 1e-11 * scale
    # device-resident output and the packed host output agree bitwise
    dev = gpu.alloc(got.size * 16)
    try:
        db.eri_derivative_batch(qs, _f6(field), device_out=dev)
        on_device = gpu.to_host(dev, np.empty(got.size, dtype=np.complex128))
    finally:
        gpu.free(dev)
    assert np.array_equal(on_device, got)


@pytest.mark.parametrize("field", FIELDS)
def test_pair_bounds_match_oracle_and_dominate(gpu, port, rng, field):
    raw = random_shells(rng, [0, 1, 2, 1], nprim=(1, 3), ncontr=(1, 3))
    b = make_basis(raw, port)
    basis = _core_basis([(c, l, e, k) for c, l, e, k in raw])
    bounds = core._EriDerivativeSchwarzBounds(basis, field=field)
    ref = port.eri_derivative_pair_bounds(b, _f6(field))
    rel = float(np.max(np.abs(bounds._values - ref) / ref))
    print(f"derivative pair bounds: max relative difference {rel:.3e}", flush=True)
    assert rel <= 1e-12
    qs = list(product(range(4), repeat=4))
    values = [core.eri_center_derivatives_shell(*[basis.shells[k] for k in q], field=field) for q in qs]
    for q, v in zip(qs, values):
        assert np.max(np.abs(v)) <= bounds.quartet_bound(q) * (1 + 1e-13)


@pytest.mark.parametrize("field", FIELDS)
@pytest.mark.parametrize("explicit", [False, True])
def test_streamed_blocks(field, explicit):
    """test_eri_derivative_batches.py:23-56 on the device."""
    basis = core.Basis([core.Shell((0.1, -0.2, 0.3), 0, [0.8, 1.2], [0.7, -0.2]),
                        core.Shell((-0.3, 0.4, 0.2), 1, [0.6], [1.0])])
    qs = list(product(range(2), repeat=4)) if explicit else None
    batches = list(eri.eri_derivative_batches(basis, quartets=qs, target_bytes=3000, field=field))
    again = list(eri.eri_derivative_batches(basis, quartets=qs, target_bytes=3000, field=field, threads=2))
    assert len(batches) == len(again) > 1
    seen = []
    for x, y in zip(batches, again):
        assert isinstance(x, eri.EriDerivativeBatch)
        assert x.shapes.shape[1] == 6 and x.offsets[-1] == x.values.size and x.screened_count == 0
        if x.values.nbytes > 3000:
            assert len(x.quartets) == 1
        for attr in ("quartets", "shapes", "offsets", "values"):
            assert np.array_equal(getattr(x, attr), getattr(y, attr))
        for row, q in enumerate(x.quartets):
            seen.append(tuple(int(i) for i in q))
            expected = core.eri_center_derivatives_shell(*[basis.shells[int(k)] for k in q], field=field)
            assert np.array_equal(x.block(row), expected)
            assert np.shares_memory(x.block(row), x.values)
    if explicit:
        assert seen == qs
    else:
        assert seen == [tuple(int(i) for i in q) for batch in eri.eri_batches(basis, field=field)
                        for q in batch.quartets]


@pytest.mark.parametrize("field", FIELDS)
def test_screening_is_safe(field):
    """test_eri_derivative_batches.py:76-121,207-240 on the device."""
    for far, threshold in (((30, -20, 10), 1e-8), ((4, 0, 0), None)):
        basis = core.Basis([core.Shell((0.1, -0.2, 0.3), 0, [0.8], [1.0]), core.Shell(far, 1, [0.9], [1.0])])
        qs = list(product(range(2), repeat=4))
        if threshold is None:
            bounds = core._EriDerivativeSchwarzBounds(basis, field=field)
            threshold = float(np.median([bounds.quartet_bound(q) for q in qs]))
        out = list(eri.eri_derivative_batches(basis, quartets=qs, field=field, target_bytes=2000,
                                              screening_threshold=threshold))
        assert sum(x.requested_count for x in out) == len(qs)
        assert 0 < sum(x.screened_count for x in out) < len(qs)
        kept = set()
        for x in out:
            for row, q in enumerate(x.quartets):
                q = tuple(int(i) for i in q)
                kept.add(q)
                assert np.array_equal(x.block(row),
                                      core.eri_center_derivatives_shell(*[basis.shells[k] for k in q], field=field))
        for q in qs:
            if q not in kept:
                full = core.eri_center_derivatives_shell(*[basis.shells[k] for k in q], field=field)
                assert np.max(np.abs(full)) < threshold


@pytest.mark.parametrize("field", FIELDS)
def test_finite_differences_and_molecule(field):
    """test_eri_derivative_batches.py:124-170: both occurrences of a shell move
    together in a shell derivative (repeated slots add)."""
    centers = np.array([[0.1, -0.2, 0.3], [-0.3, 0.4, 0.2]])

    def shells(c):
        return [core.Shell(c[0], 0, [0.8], [1.0]), core.Shell(c[1], 1, [0.6], [1.0])]

    q = (0, 1, 0, 1)
    block = next(eri.eri_derivative_batches(core.Basis(shells(centers)), quartets=[q], field=field)).block(0)
    for center in range(2):
        for axis in range(3):
            def value(step):
                c = centers.copy()
                c[center, axis] += step
                s = shells(c)
                return core.eri_shell(*[s[k] for k in q], field=field)

            step = 4e-4
            coarse = (value(step) - value(-step)) / (2 * step)
            fine = (value(step / 2) - value(-step / 2)) / step
            expected = (4 * fine - coarse) / 3
            actual = block[[k for k, s in enumerate(q) if s == center], axis].sum(axis=0)
            np.testing.assert_allclose(actual, expected, atol=3e-10, rtol=3e-9)
    mol = Molecule([("H", (0.1, -0.2, 0.3)), ("H", (-0.3, 0.4, 0.2))], workloads.DATA / "cc-pvdz-h-o.json",
                   unit="bohr", field=field)
    got = next(mol.eri_derivative_batches(quartets=[q]))
    want = next(eri.eri_derivative_batches(mol.basis, quartets=[q], field=field))
    assert np.array_equal(got.values, want.values)


@pytest.mark.parametrize("field", FIELDS)
def test_water_ccpvtz_sum_rule(field, rng):
    """A size-independent identity on 300 quartets of water in cc-pVTZ (general
    contractions, up to f): translating all four centres multiplies (ab|cd)
    by exp(i/2 (B x (A-B) + B x (C-D)) . d)."""
    mol = Molecule(WATER_BOHR, workloads.DATA / "cc-pvtz-h-o.json", unit="bohr", field=field)
    basis = mol.basis
    ns = len(basis.shells)
    qs = [tuple(int(i) for i in rng.integers(0, ns, 4)) for _ in range(300)]
    q, shapes, offsets, values, _, _ = core._eri_derivative_batch(basis, qs, field=field)
    kq, kshapes, koffsets, kvalues, _, _ = core._eri_batch(basis, qs, field=field)
    B = _f6(field)[:3]
    worst = scale = 0.0
    for i, (a, b, c, d) in enumerate(qs):
        D = values[offsets[i]:offsets[i + 1]].reshape(4, 3, -1)
        V = kvalues[koffsets[i]:koffsets[i + 1]]
        A = [np.asarray(basis.shells[k].center) for k in (a, b, c, d)]
        phase = 0.5j * (np.cross(B, A[0] - A[1]) + np.cross(B, A[2] - A[3]))
        worst = max(worst, float(np.max(np.abs(D.sum(axis=0) - phase[:, None] * V[None, :]))))
        scale = max(scale, float(np.max(np.abs(D))))
    print(f"water cc-pVTZ sum rule over {len(qs)} quartets: max residual {worst:.3e}, scale {scale:.3e}", flush=True)
    assert worst <= 1e-12 * scale


def test_errors_and_empty():
    """test_eri_derivative_batches.py:173-197 and the binding's checks."""
    basis = core.Basis([core.Shell((0.1, -0.2, 0.3), 0, [0.8, 1.2], [0.7, -0.2]),
                        core.Shell((-0.3, 0.4, 0.2), 1, [0.6], [1.0])])
    for bad in (0, -1, True, 1.5):
        with pytest.raises(ValueError):
            list(eri.eri_derivative_batches(basis, target_bytes=bad))
    for bad in (-1, float("nan"), float("inf")):
        with pytest.raises(ValueError):
            list(eri.eri_derivative_batches(basis, screening_threshold=bad))
    assert list(eri.eri_derivative_batches(core.Basis([]))) == []
    assert list(eri.eri_derivative_batches(basis, quartets=[])) == []
    for q in [(0, 0, 0, 9), (0, 0, 0, True), (0, 0, 0, 0.5)]:
        with pytest.raises((ValueError, IndexError)):
            list(eri.eri_derivative_batches(basis, quartets=[q]))
    with pytest.raises(IndexError, match="derivative shell quartet index is out of range"):
        core._eri_derivative_batch(basis, [(0, 0, 0, 2)])
    with pytest.raises(ValueError, match="derivative screening requires derivative bounds"):
        core._eri_derivative_batch(basis, [(0, 0, 0, 1)], screening_threshold=1e-8)
    bounds = core._EriDerivativeSchwarzBounds(basis)
    with pytest.raises(ValueError, match="derivative bounds do not match basis and field"):
        core._eri_derivative_batch(basis, [(0, 0, 0, 1)], field=FIELDS[1], bounds=bounds)
    # (8 8|8 8): the shifted classes pass the reference's combined order 32
    # (its test uses an l = 33 shell, test_eri_derivative_batches.py:186-197)
    b8 = core.Basis([core.Shell((0, 0, 0.1 * k), 8, [0.6], [1.0]) for k in range(4)])
    with pytest.raises(ValueError, match="angular order through 32"):
        list(eri.eri_derivative_batches(b8, quartets=[(0, 1, 2, 3)]))
    # one more unit of angular momentum than the ERI kernels admit
    high = core.Basis([core.Shell((0, 0, 0), 0, [0.8], [1.0]), core.Shell((0, 0, 0), 8, [0.6], [1.0])])
    with pytest.raises(RuntimeError, match="angular momentum through 8"):
        list(eri.eri_derivative_batches(high, quartets=[(1, 0, 0, 0)]))
