"""Molecule front end (python/giao_integrals/molecule.py:26-450, ERI part)."""

import json

import numpy as np
import pytest

from giaoint_amd import core, workloads
from giaoint_amd.molecule import BOHR_IN_ANGSTROM, Molecule

WATER = [("O", (0.0, 0.0, 0.1173)), ("H", (0.0, 0.7572, -0.4692)), ("h", (0.0, -0.7572, -0.4692))]
BASIS = workloads.DATA / "cc-pvdz-h-o.json"


def test_water_basis_order_and_units():
    mol = Molecule(WATER, BASIS, field=core.MagneticField((0, 0, 0.1)))
    assert mol.ao_count == 25
    assert [s for s, _ in mol.atoms] == ["O", "H", "H"]
    assert mol.atoms[1][1][1] == pytest.approx(0.7572 / BOHR_IN_ANGSTROM, rel=1e-15)
    # atom order, then stable increasing L; general contractions kept as one shell
    ls = [s.angular_momentum for s in mol.basis.shells]
    assert ls == [0, 1, 2, 0, 1, 0, 1]
    assert mol.shell_atoms == (0, 0, 0, 1, 1, 2, 2)
    ref, _, _ = workloads.water_ccpvdz()
    assert [s.contraction_count for s in mol.basis.shells] == list(ref.ncontr)
    assert [s.primitive_count for s in mol.basis.shells] == list(ref.nprim)
    np.testing.assert_allclose(np.array([s.center for s in mol.basis.shells]), ref.centers, rtol=0, atol=1e-14)
    bohr = Molecule([(s, tuple(x / BOHR_IN_ANGSTROM for x in p)) for s, p in WATER], json.loads(BASIS.read_text()),
                    unit="bohr")
    np.testing.assert_allclose(np.array([p for _, p in bohr.atoms]), np.array([p for _, p in mol.atoms]), atol=1e-14)


def test_xyz_and_errors():
    text = "3\nwater\nO 0 0 0.1173\nH 0 0.7572 -0.4692\n1 0 -0.7572 -0.4692\n"
    mol = Molecule.from_xyz(text, BASIS)
    assert [s for s, _ in mol.atoms] == ["O", "H", "H"]
    with pytest.raises(ValueError, match="unit must be 'angstrom' or 'bohr'"):
        Molecule(WATER, BASIS, unit="nm")
    with pytest.raises(ValueError, match="Unknown element"):
        Molecule([("Xx", (0, 0, 0))], BASIS)
    with pytest.raises(ValueError, match="Basis has no data for element C"):
        Molecule([("C", (0, 0, 0))], BASIS)
    with pytest.raises(ValueError, match="at least one atom"):
        Molecule([], BASIS)
    with pytest.raises(ValueError, match="XYZ requires a comment line"):
        Molecule.from_xyz("2\nx\nH 0 0 0\n", BASIS)
    with pytest.raises(ImportError):
        Molecule(WATER, "cc-pvdz")  # named bases need basis-set-exchange
    with pytest.raises(TypeError):
        Molecule(WATER, BASIS, field=(0, 0, 1))
