"""Device-resident strict-pivot Cholesky against the checker driver.

The checker (oracle/cholesky_driver.py, a restatement of
python/giao_integrals/cholesky.py:60-297) runs on the C oracle's diagonal
and columns. Pivots must agree except where the two candidates' residuals
tie to 1e-12 relative (a rounding-level tie may reorder them); factors must
agree to 1e-10 up to that point; the final residual must be below the
stopping tolerance. Exact ties are common: the ordered pairs (mu nu) and
(nu mu) have equal diagonals, so the pivot order within such a pair is a
rounding decision; the tests bound the number of tie decisions taken the
other way (at most half the rank) and print it.
"""

import numpy as np
import pytest

from giaoint_amd import _native, workloads
from oracle import cholesky_driver
from oracle.pyoracle import make_basis

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("natoms", [2, 3])
def test_device_cholesky_matches_checker(gpu, port, natoms):
    """Identical even-tempered atoms make many pair functions exactly equal
    by symmetry, so the first maximum is decided by rounding. The checker
    therefore replays the GPU's pivots and requires each to be a maximum up
    to 1e-12 * max(initial diagonal); every factor must then agree to 1e-10."""
    sh, B, O = workloads.cholesky_atoms(natoms=natoms)
    b = make_basis(sh.shell_list(), port)
    f = _native.field6(B, O)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    res = _native.cholesky_run(db, f, 1e-8)
    assert res.converged
    assert res.max_residual <= res.tolerance
    gp = res.pivots()
    stats = {}
    vec, piv, maxres, tol = cholesky_driver.cholesky_eri(
        b.ao_offsets, [((l + 1) * (l + 2) // 2, 1, 1) for l in b.L],
        lambda pairs: port.cholesky_diagonal(b, pairs, f, 16),
        lambda pv: port.cholesky_columns(b, pv, f, 16),
        real=False, single_center=False, threshold=1e-8, follow=gp, stats=stats,
    )
    print(f"rank {res.rank}; pivots decided by rounding-level ties: {stats.get('tie_flips', 0)}")
    assert stats.get("tie_flips", 0) <= res.rank // 2
    assert abs(res.tolerance - tol) <= 1e-12 * tol
    assert len(piv) == res.rank
    worst = 0.0
    for first in range(0, res.rank, 256):
        cnt = min(256, res.rank - first)
        worst = max(worst, float(np.max(np.abs(res.vectors(first, cnt) - vec[first:first + cnt]))))
    print(f"max |L_gpu - L_checker| = {worst:.3e}")
    assert worst <= 1e-10
    assert abs(res.max_residual - maxres) <= 1e-12


def test_device_cholesky_rank_cap_and_prefix(gpu, port):
    sh, B, O = workloads.cholesky_atoms(natoms=2)
    b = make_basis(sh.shell_list(), port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    f = _native.field6(B, O)
    with pytest.raises(MemoryError):
        _native.cholesky_run(db, f, 1e-8, max_rank=5)
    pre = _native.cholesky_run(db, f, 1e-8, stop_after=7)
    assert pre.rank == 7 and not pre.converged


def test_device_cholesky_resume_matches_single_run(gpu, port):
    """A run stopped after a prefix and continued in steps from its device
    state gives exactly the pivots and factors of an uninterrupted run."""
    sh, B, O = workloads.cholesky_atoms(natoms=2)
    b = make_basis(sh.shell_list(), port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    f = _native.field6(B, O)
    full = _native.cholesky_run(db, f, 1e-8)
    part = _native.cholesky_run(db, f, 1e-8, stop_after=7)
    assert part.rank == 7 and not part.converged
    target = 7
    while not part.converged:
        target += 25
        part.resume(target)
        assert part.rank == min(target, full.rank)
    part.resume(target + 10)  # a converged run stays as it is
    assert part.rank == full.rank and part.converged
    np.testing.assert_array_equal(part.pivots(), full.pivots())
    np.testing.assert_array_equal(part.vectors(0, part.rank), full.vectors(0, full.rank))
    assert part.max_residual == full.max_residual


def _run_with_block(ctx, db, f, nb, **kw):
    ctx.set_cholesky_block(nb)
    try:
        return _native.cholesky_run(db, f, 1e-8, **kw)
    finally:
        ctx.set_cholesky_block(32)


def test_device_cholesky_block_sizes_bitwise(gpu, port):
    """Pivot blocks only decide which candidate projections share a pass over
    the factors; every projection sums the same terms in the same order.
    Runs with 1, 3, 16 and 32 candidates per block agree bitwise."""
    sh, B, O = workloads.cholesky_atoms(natoms=3)
    b = make_basis(sh.shell_list(), port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    f = _native.field6(B, O)
    ref = _run_with_block(gpu, db, f, 1)
    assert ref.converged
    # one candidate per block: one pass per pivot over the earlier factors in
    # whole groups of four rows (the catch-up adds the last rank % 4 rows)
    assert ref.blocks["blocks"] == ref.rank
    assert ref.blocks["factor_rows_read"] == sum(r - r % 4 for r in range(ref.rank))
    assert ref.blocks["candidate_rows"] == ref.blocks["factor_rows_read"]
    ref_piv, ref_vec = ref.pivots(), ref.vectors()
    for nb in (3, 16, 32):
        got = _run_with_block(gpu, db, f, nb)
        print(f"block {nb}: rank {got.rank}, {got.blocks['blocks']} blocks")
        assert got.blocks["blocks"] < ref.blocks["blocks"]
        # every candidate of a block is projected against the rank at the block start
        assert got.blocks["candidate_rows"] <= got.blocks["projected"] * got.blocks["factor_rows_read"]
        assert got.blocks["candidate_rows"] >= got.blocks["factor_rows_read"]
        assert got.rank == ref.rank and got.converged
        np.testing.assert_array_equal(got.pivots(), ref_piv)
        np.testing.assert_array_equal(got.vectors(), ref_vec)
        assert got.max_residual == ref.max_residual
    for bad in (0, 33):
        with pytest.raises(ValueError):
            gpu.set_cholesky_block(bad)


@pytest.mark.parametrize("natoms,zero_field,stop", [(3, False, 0), (3, True, 0), (10, False, 900)])
def test_device_cholesky_graphs_bitwise(gpu, port, natoms, zero_field, stop):
    """The column launches of each pivot as hipGraphs (one executable graph
    per launch topology, updated from every pivot's captured launches and
    replayed between the pivot's own direct launches) give the pivots and
    factors of direct launches bitwise, in a field and at zero field (real
    columns)."""
    sh, B, O = workloads.cholesky_atoms(natoms=natoms)
    if zero_field:
        B = (0.0, 0.0, 0.0)
    b = make_basis(sh.shell_list(), port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    f = _native.field6(B, O)
    runs = {}
    try:
        for on in (False, True, "inside"):
            print(f"{natoms} atoms, B={B}, graphs {on}, stop at {stop}", flush=True)
            gpu.set_cholesky_graphs(on)
            runs[on] = _native.cholesky_run(db, f, 1e-8, stop_after=stop)
    finally:
        gpu.set_cholesky_graphs(False)
    direct = runs[False]
    assert direct.graphs["launches"] == 0
    for mode in (True, "inside"):
        graphs = runs[mode]
        print(f"graphs {mode}: rank {graphs.rank}, graph use {graphs.graphs}")
        g = graphs.graphs
        # every pivot after the first of its topology replays a graph
        assert g["launches"] > graphs.rank // 2
        assert g["instantiated"] + g["updated"] == g["launches"]
        assert g["instantiated"] < g["launches"]
        assert graphs.rank == direct.rank and graphs.converged == direct.converged
        np.testing.assert_array_equal(graphs.pivots(), direct.pivots())
        np.testing.assert_array_equal(graphs.vectors(), direct.vectors())
        assert graphs.max_residual == direct.max_residual


def test_device_cholesky_blocks_across_factor_chunks(gpu, port):
    """Ten atoms (350 AOs): the factors span two device chunks of 1095 rows,
    so projections cross a chunk boundary. Blocked (16) and one-at-a-time
    runs stopped at rank 1400 agree bitwise."""
    sh, B, O = workloads.cholesky_atoms(natoms=10)
    b = make_basis(sh.shell_list(), port)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    f = _native.field6(B, O)
    one = _run_with_block(gpu, db, f, 1, stop_after=1400)
    blk = _run_with_block(gpu, db, f, 16, stop_after=1400)
    assert one.rank == blk.rank == 1400
    np.testing.assert_array_equal(blk.pivots(), one.pivots())
    for first in range(0, 1400, 200):
        np.testing.assert_array_equal(blk.vectors(first, 200), one.vectors(first, 200))
    assert blk.max_residual == one.max_residual


@pytest.mark.parametrize("case", ["zero-field", "one-centre"])
def test_device_cholesky_real_arithmetic(gpu, port, case):
    """Zero-field and one-centre bases run in real arithmetic with
    symmetrized columns (cholesky.py:128-133,175-176,269-271); on one centre
    the integrals are taken at zero field whatever B is. Checker replay as
    for the complex runs; the factors are real and symmetric."""
    if case == "zero-field":
        sh, _, _ = workloads.cholesky_atoms(natoms=2)
        B, O = (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)
    else:
        sh, B, O = workloads.cholesky_atoms(natoms=1)
    b = make_basis(sh.shell_list(), port)
    f = _native.field6(B, O)
    fint = f if case == "zero-field" else _native.field6()
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    res = _native.cholesky_run(db, f, 1e-8)
    assert res.converged and res.max_residual <= res.tolerance
    gp = res.pivots()
    stats = {}
    vec, piv, maxres, tol = cholesky_driver.cholesky_eri(
        b.ao_offsets, [((l + 1) * (l + 2) // 2, 1, 1) for l in b.L],
        lambda pairs: port.cholesky_diagonal(b, pairs, fint, 16),
        lambda pv: port.cholesky_columns(b, pv, fint, 16, real=True),
        real=True, single_center=(case == "one-centre"), threshold=1e-8, follow=gp, stats=stats,
    )
    print(f"{case}: rank {res.rank}; rounding-level tie decisions {stats.get('tie_flips', 0)}")
    assert stats.get("tie_flips", 0) <= res.rank // 2
    assert len(piv) == res.rank
    got = res.vectors()
    assert np.all(got.imag == 0.0)
    np.testing.assert_array_equal(got, got.transpose(0, 2, 1))
    assert np.max(np.abs(got - vec)) <= 1e-10
    assert abs(res.max_residual - maxres) <= 1e-12


def test_device_factor_consumers(gpu, port):
    """Reconstruction of AO blocks and the Coulomb contraction
    J[mu,nu] = sum (mu nu|lam sig) D[lam,sig] computed from the device factors,
    against numpy on the fetched factors and against the ERI tensor itself
    (reconstruction bound: every element within the threshold)."""
    sh, B, O = workloads.cholesky_atoms(natoms=2)
    b = make_basis(sh.shell_list(), port)
    f = _native.field6(B, O)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    res = _native.cholesky_run(db, f, 1e-8)
    L = res.vectors()
    n = db.nao
    got = res.reconstruct((3, 20), (0, 15), (5, 30), (10, 25))
    want = np.einsum("Pmn,Psl->mnls", L[:, 3:20, 0:15], L[:, 10:25, 5:30].conj())
    assert got.shape == (17, 15, 25, 15)
    assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.abs(want).max())
    exact = db.eri_tensor(f)
    full = res.reconstruct()
    assert np.max(np.abs(full - exact)) <= 1e-8 + 1e-12
    rng = np.random.default_rng(7)
    D = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    J = res.coulomb(D)
    gamma = np.einsum("Psl,ls->P", L.conj(), D)
    Jw = np.einsum("Pmn,P->mn", L, gamma)
    assert np.max(np.abs(J - Jw)) <= 1e-11 * np.abs(Jw).max()
    Jx = np.einsum("mnls,ls->mn", exact, D)
    assert np.max(np.abs(J - Jx)) <= (1e-8 + 1e-12) * np.abs(D).sum()
    # exchange: K[mu,nu] = sum (mu lam|sig nu) D[lam,sig] ~ sum_P L_P D L_P^H
    K = res.exchange(D)
    Kw = np.einsum("Pml,ls,Pns->mn", L, D, L.conj())
    assert np.max(np.abs(K - Kw)) <= 1e-11 * np.abs(Kw).max()
    Kx = np.einsum("mlsn,ls->mn", exact, D)
    assert np.max(np.abs(K - Kx)) <= (1e-8 + 1e-12) * np.abs(D).sum()
    with pytest.raises(IndexError):
        res.reconstruct((0, n + 1))


@pytest.mark.parametrize("case", ["water", "water-bent"])
def test_device_cholesky_general_contractions(gpu, port, case):
    """Complex runs on bases with general contractions follow the reference's
    candidate-block policy (cholesky.py:193-262): the pivot may be the
    largest residual of the latest block of contraction variants instead of
    the global maximum. Water cc-pVDZ in a field (and a copy with one H moved
    off the mirror plane): the checker replays the GPU's pivots, each of
    which must be its own policy's choice up to rounding-level ties, and the
    factors must agree to 1e-10."""
    sh, B, O = workloads.water_ccpvdz()
    shells = sh.shell_list()
    if case == "water-bent":
        moved = tuple(np.add(shells[-1][0], (0.13, -0.07, 0.05)))
        shells = [s if i < len(shells) - 2 else (moved, s[1], s[2], s[3]) for i, s in enumerate(shells)]
        B = (0.03, -0.05, 0.11)
    b = make_basis(shells, port)
    assert any(k > 1 for k in b.ncontr)
    f = _native.field6(B, O)
    db = _native.DeviceBasis(gpu, b.centers, b.L, b.nprim, b.ncontr, b.exps, b.coefs)
    gpu.set_cholesky_pivot_batch(32)
    res = _native.cholesky_run(db, f, 1e-8)
    assert res.converged and res.max_residual <= res.tolerance
    gp = res.pivots()
    stats = {}
    vec, piv, maxres, tol = cholesky_driver.cholesky_eri(
        b.ao_offsets, [((l + 1) * (l + 2) // 2, k, p) for l, k, p in zip(b.L, b.ncontr, b.nprim)],
        lambda pairs: port.cholesky_diagonal(b, pairs, f, 16),
        lambda pv: port.cholesky_columns(b, pv, f, 16),
        real=False, single_center=False, threshold=1e-8, follow=gp, stats=stats,
    )
    print(f"{case}: rank {res.rank}; {stats.get('overrides', 0)} pivots from the block instead of the global "
          f"maximum; pivots decided by rounding-level ties {stats.get('tie_flips', 0)}")
    assert stats.get("tie_flips", 0) <= res.rank // 2
    assert stats.get("overrides", 0) > 0  # the policy's local override is exercised
    assert len(piv) == res.rank
    assert abs(res.tolerance - tol) <= 1e-12 * tol
    worst = float(np.max(np.abs(res.vectors(0, res.rank) - vec)))
    print(f"max |L_gpu - L_checker| = {worst:.3e}")
    assert worst <= 1e-10
    assert abs(res.max_residual - maxres) <= 1e-12
    gpu.set_cholesky_pivot_batch(32)
