"""GPU nuclear attraction (giaohip_nuclear_attraction, replacing
compute_nuclear_attraction_matrix, nuclear.cpp:282-312) against the oracle:
water cc-pVDZ at B=(0,0,0.1) a.u. and random contracted shells through
(g, g) in a field with two nuclei, at 1e-11 per element; Hermiticity; the
error paths (shells above l = 4, non-finite charges)."""

import numpy as np
import pytest

from giaoint_amd import _native, workloads
from oracle.pyoracle import make_basis

pytestmark = pytest.mark.gpu

ATOL = 1e-11


def _device(ctx, basis):
    return _native.DeviceBasis(ctx, basis.centers, basis.L, basis.nprim, basis.ncontr, basis.exps, basis.coefs)


def test_nuclear_water(gpu, port):
    sh, B, O = workloads.water_ccpvdz()
    basis = make_basis(sh.shell_list(), port)
    centers = np.array([s[0] for s in sh.shell_list()])
    _, idx = np.unique(centers.round(12), axis=0, return_index=True)
    nuc = centers[np.sort(idx)]
    z = np.array([8.0 if abs(c[1]) < 1e-9 else 1.0 for c in nuc])
    f = _native.field6(B, O)
    print("nuclear attraction: water cc-pVDZ, 3 nuclei", flush=True)
    got = _device(gpu, basis).nuclear_attraction(z, nuc, f)
    want = port.nuclear_matrix(basis, z, nuc, f)
    err = np.max(np.abs(got - want))
    print(f"  max |GPU - oracle| = {err:.2e}, max |V| = {np.max(np.abs(want)):.2e}")
    assert err <= ATOL
    assert np.max(np.abs(got - got.conj().T)) <= ATOL


def test_nuclear_random_shells_through_g(gpu, port, rng):
    from tests.conftest import random_shells

    shells = random_shells(rng, [4, 0, 1, 2, 3, 4, 2], nprim=(1, 4), ncontr=(1, 3))
    basis = make_basis(shells, port)
    z = np.array([3.0, 1.0])
    c = rng.uniform(-1.5, 1.5, (2, 3))
    f = _native.field6((0.2, -0.3, 0.4), (0.1, 0.0, -0.2))
    print("nuclear attraction: 7 random shells through (g, g), 2 nuclei", flush=True)
    got = _device(gpu, basis).nuclear_attraction(z, c, f)
    want = port.nuclear_matrix(basis, z, c, f)
    err = np.max(np.abs(got - want))
    print(f"  max |GPU - oracle| = {err:.2e}, max |V| = {np.max(np.abs(want)):.2e}")
    assert err <= ATOL


def test_nuclear_error_paths(gpu, port):
    basis = make_basis([(np.zeros(3), 5, np.array([1.0]), np.array([[1.0]]))], port)
    db = _device(gpu, basis)
    with pytest.raises(RuntimeError, match="through l = 4"):
        db.nuclear_attraction([1.0], [[0.0, 0.0, 1.0]], _native.field6())
    small = _device(gpu, make_basis([(np.zeros(3), 1, np.array([1.0]), np.array([[1.0]]))], port))
    with pytest.raises(ValueError, match="finite"):
        small.nuclear_attraction([np.nan], [[0.0, 0.0, 1.0]], _native.field6())


def test_nuclear_reference_named_entry(port):
    """giaoint_amd.core.nuclear_attraction (the pybind11 name) with Nucleus
    objects and a strict out= array gives the C-ABI matrix."""
    from giaoint_amd import core

    sh, B, O = workloads.water_ccpvdz()
    shells = sh.shell_list()
    basis = core.Basis([core.Shell(c, l, e, k) for c, l, e, k in shells])
    nuclei = [core.Nucleus(8.0, shells[0][0])] + [core.Nucleus(1.0, c) for c in
                                                   {tuple(s[0]) for s in shells if abs(s[0][1]) > 1e-9}]
    field = core.MagneticField(B, O)
    out = np.empty((basis.ao_count, basis.ao_count), dtype=np.complex128)
    got = core.nuclear_attraction(basis, nuclei, field=field, out=out)
    assert got is out
    z = np.array([n.charge for n in nuclei])
    c = np.array([n.center for n in nuclei])
    want = port.nuclear_matrix(make_basis(shells, port), z, c, _native.field6(B, O))
    assert np.max(np.abs(out - want)) <= ATOL
    with pytest.raises(ValueError):
        core.nuclear_attraction(basis, nuclei, field=field, out=np.empty((2, 2), dtype=np.complex128))
