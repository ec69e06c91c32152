"""Host-side behaviour of the reference-named Python layer (giaoint_amd.core,
.eri, .cholesky) that needs no GPU: argument checks and their messages
(bindings.cpp:37-191, eri.py:61-201, cholesky.py:98-125, _threads.py:21-30),
shell bookkeeping, the London wave vector and quartet canonicalization."""

import itertools

import numpy as np
import pytest

from giaoint_amd import cholesky, core, eri


def test_field_and_wave_vector():
    f = core.MagneticField((0.1, -0.2, 0.3), (1.0, 0.0, -1.0))
    B, O, C = np.array(f.B), np.array(f.gauge_origin), np.array([0.5, 2.0, -3.0])
    np.testing.assert_allclose(f.london_wave_vector(C), 0.5 * np.cross(B, C - O), rtol=0, atol=1e-15)
    assert core.MagneticField().london_wave_vector((1, 2, 3)) == (0.0, 0.0, 0.0)
    with pytest.raises(ValueError, match="one-dimensional array with length 3"):
        core.MagneticField((0.0, 1.0))


def test_shell_bookkeeping_and_checks():
    s = core.Shell((0, 0, 0), 2, [1.0, 0.3], [[0.5, 0.5], [1.0, -1.0]])
    assert (s.primitive_count, s.contraction_count, s.cartesian_count, s.ao_count) == (2, 2, 6, 12)
    assert s.components[:3] == ((2, 0, 0), (1, 1, 0), (1, 0, 1))
    raw = core.Shell((0, 0, 0), 1, [0.7], [1.0], normalize=False)
    assert raw._coefs.tolist() == [[1.0]]
    with pytest.raises(ValueError, match="final coefficient dimension"):
        core.Shell((0, 0, 0), 0, [1.0, 2.0], [1.0])
    with pytest.raises(ValueError, match="one- or two-dimensional"):
        core.Shell((0, 0, 0), 0, [1.0], np.ones((1, 1, 1)))
    with pytest.raises(ValueError):
        core.Shell((0, 0, 0), 0, [-1.0], [1.0])
    b = core.Basis([s, raw])
    assert b.ao_offsets == (0, 12) and b.ao_count == 15
    with pytest.raises(ValueError, match="exactly three integers"):
        core.PrimitiveGaussian(1.0, (0, 0, 0), angular=(1, 0))
    with pytest.raises(ValueError, match="finite and positive"):
        core.PrimitiveGaussian(0.0, (0, 0, 0))


def test_canonical_quartet_matches_oracle(port):
    import ctypes as C

    fn = port._fn("canonicalize")
    for q in itertools.product(range(3), repeat=4):
        out = (C.c_uint32 * 4)()
        conj = fn((C.c_uint32 * 4)(*q), out)
        got, flag = core._canonical_shell_quartet(*q)
        assert got == tuple(out) and flag == bool(conj)


def test_threads_and_threshold_checks():
    assert core._threads(None) == 1 and core._threads(8) == 8
    for bad in (0, -1, True, 1.5, 2**31):
        with pytest.raises(ValueError, match="positive integer"):
            core._threads(bad)
    s = core.Shell((0, 0, 0), 0, [1.0], [1.0])
    b = core.Basis([s])
    with pytest.raises(ValueError, match="storage must be"):
        eri.eri(b, storage="dense")
    with pytest.raises(ValueError, match="requires an explicit max_bytes"):
        eri.eri(b, storage="full")
    with pytest.raises(ValueError, match="only valid with storage='full'"):
        eri.eri(b, max_bytes=10)
    with pytest.raises(MemoryError, match="exceeding max_bytes=15"):
        eri.eri(b, storage="full", max_bytes=15)
    with pytest.raises(ValueError, match="finite and non-negative"):
        eri.eri(b, screening_threshold=-1.0)
    with pytest.raises(ValueError, match="target_bytes"):
        next(eri.eri_batches(b, target_bytes=0))
    with pytest.raises(IndexError, match="out of range"):
        next(eri.eri_batches(b, quartets=[(0, 0, 0, 1)]))
    with pytest.raises(ValueError, match="four indices"):
        next(eri.eri_batches(b, quartets=[(0, 0, 0)]))
    with pytest.raises(ValueError, match="must be integers"):
        next(eri.eri_batches(b, quartets=[(0, 0, 0, 0.5)]))
    canon = list(eri._quartets(3, None))
    assert len(canon) == (3**4 + 3 * 3**2) // 4


def test_cholesky_argument_checks():
    s1 = core.Shell((0, 0, 0), 0, [1.0], [1.0])
    s2 = core.Shell((0, 0, 1.5), 0, [1.0], [1.0])
    b = core.Basis([s1, s2])
    f = core.MagneticField((0.0, 0.0, 0.1))
    with pytest.raises(ValueError, match="threshold must be finite and non-negative"):
        cholesky.cholesky_eri(b, field=f, threshold=-1.0)
    with pytest.raises(ValueError, match="pivot_batch_size"):
        cholesky.cholesky_eri(b, field=f, pivot_batch_size=0)
    with pytest.raises(ValueError, match="max_bytes"):
        cholesky.cholesky_eri(b, field=f, max_bytes=-1)


def test_reconstruct_convention():
    rng = np.random.default_rng(3)
    v = rng.normal(size=(4, 3, 3)) + 1j * rng.normal(size=(4, 3, 3))
    r = cholesky.CholeskyERI(v, np.zeros((4, 2), dtype=np.int64), 1e-8, 0.0)
    want = np.einsum("Pmn,Psl->mnls", v, v.conj())
    np.testing.assert_allclose(r.reconstruct(), want, atol=1e-13)
    sl = (slice(0, 2), slice(1, 3), slice(0, 3), slice(2, 3))
    np.testing.assert_allclose(r.reconstruct(sl), want[0:2, 1:3, 0:3, 2:3], atol=1e-13)
    with pytest.raises(ValueError, match="four slices"):
        r.reconstruct((slice(None),) * 3)


def test_nucleus_checks():
    """Nucleus rejects what types.cpp:188-194 rejects, with its messages."""
    for charge in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(ValueError, match="nuclear charge must be finite and positive"):
            core.Nucleus(charge, (0.0, 0.0, 0.0))
    with pytest.raises(ValueError, match="nuclear center values must be finite"):
        core.Nucleus(8.0, (0.0, float("nan"), 0.0))
    with pytest.raises(ValueError, match="center must be a one-dimensional array with length 3"):
        core.Nucleus(8.0, (0.0, 0.0))
    nu = core.Nucleus(1, (0.0, 0.5, -1.0))
    assert nu.charge == 1.0 and nu.center == (0.0, 0.5, -1.0)


def test_nuclear_attraction_checks_out_before_any_work():
    """A bad out= is rejected before the device is touched (bindings.cpp:554-573
    validates the output first), so this needs no GPU."""
    b = core.Basis([core.Shell((0, 0, 0), 1, [1.0], [1.0])])
    with pytest.raises(ValueError, match="out has an incorrect shape"):
        core.nuclear_attraction(b, [core.Nucleus(1.0, (0, 0, 0))], out=np.zeros((2, 2), np.complex128))
    with pytest.raises(ValueError, match="out must have dtype complex128"):
        core.nuclear_attraction(b, [core.Nucleus(1.0, (0, 0, 0))], out=np.zeros((3, 3)))


def test_openmp_max_threads_reports_library_threads():
    assert 1 <= core.openmp_max_threads() <= 16
