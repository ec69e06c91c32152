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


def test_xyz_details(tmp_path):
    """molecule.py:86-112: BOM, Fortran exponents, blank lines, one frame, line numbers."""
    text = "\ufeff3\nwater\nO 0 0 1.173D-1\nH 0 0.7572 -0.4692\nH 0 -0.7572 -0.4692\n\n"
    mol = Molecule.from_xyz(text, BASIS)
    assert mol.atoms[0][1][2] == pytest.approx(0.1173 / BOHR_IN_ANGSTROM, rel=1e-15)
    path = tmp_path / "w.xyz"
    path.write_text(text, encoding="utf-8")
    assert Molecule.from_xyz(path, BASIS).atoms == mol.atoms
    assert Molecule.from_xyz(str(path), BASIS).atoms == mol.atoms
    assert Molecule.from_xyz("\n\n" + text.lstrip("\ufeff"), BASIS).atoms == mol.atoms
    with pytest.raises(ValueError, match="integer atom count"):
        Molecule.from_xyz("three\nx\nH 0 0 0\n", BASIS)
    with pytest.raises(ValueError, match="atom count must be positive"):
        Molecule.from_xyz("0\nx\n", BASIS)
    with pytest.raises(ValueError, match="Expected one XYZ frame"):
        Molecule.from_xyz("1\nx\nH 0 0 0\n1\ny\nH 0 0 1\n", BASIS)
    with pytest.raises(ValueError, match="XYZ line 4 must contain an element and three coordinates"):
        Molecule.from_xyz("2\nx\nH 0 0 0\nH 0 0\n", BASIS)
    with pytest.raises(ValueError, match=r"Invalid atom 2: coordinates must be three finite numbers"):
        Molecule([("H", (0, 0, 0)), ("H", (0, 0, float("nan")))], BASIS)
    with pytest.raises(ValueError, match=r"Invalid atom 1: "):
        Molecule([("H", (0, 0))], BASIS)


def test_basis_entries_combined_shells_and_errors():
    """Combined (SP) shells split per L with their own row; general
    contractions stay one shell; stable L order within an element; errors
    name the element and entry (molecule.py:191-267)."""
    data = {"elements": {"1": {"electron_shells": [
        {"function_type": "gto", "angular_momentum": [1], "exponents": ["0.8"], "coefficients": [["1.0"]]},
        {"function_type": "gto", "angular_momentum": [0, 1], "exponents": ["3.0", "0.5"],
         "coefficients": [["0.4", "0.7"], ["0.2", "0.9"]]},
        {"function_type": "gto", "angular_momentum": [0], "exponents": ["1.0", "0.2"],
         "coefficients": [["1.0", "0.0"], ["0.0", "1.0"]]},
    ]}}}
    mol = Molecule([("H", (0, 0, 0)), ("H", (0, 0, 1.4))], data, unit="bohr")
    ls = [s.angular_momentum for s in mol.basis.shells]
    assert ls == [0, 0, 1, 1] * 2
    assert [s.contraction_count for s in mol.basis.shells[:4]] == [1, 2, 1, 1]
    assert [s.primitive_count for s in mol.basis.shells[:4]] == [2, 2, 1, 2]
    assert mol.shell_atoms == (0, 0, 0, 0, 1, 1, 1, 1)
    assert mol.basis.shells[4].center == (0.0, 0.0, 1.4)
    bad = {"elements": {"1": {"electron_shells": [{"angular_momentum": [0, 0], "exponents": [1.0],
                                                   "coefficients": [[1.0], [1.0]]}]}}}
    with pytest.raises(ValueError, match="Invalid basis shell 0 for H: combined shells must have distinct"):
        Molecule([("H", (0, 0, 0))], bad)
    ecp = {"elements": {"1": {"ecp_electrons": 2, "electron_shells": [{}]}}}
    with pytest.raises(ValueError, match="contains an ECP"):
        Molecule([("H", (0, 0, 0))], ecp)
    with pytest.raises(ValueError, match="has no electron shells"):
        Molecule([("H", (0, 0, 0))], {"elements": {"1": {"electron_shells": []}}})
    with pytest.raises(ValueError, match="'elements' object"):
        Molecule([("H", (0, 0, 0))], {"name": "x"})
    sto = {"elements": {"1": {"electron_shells": [{"angular_momentum": [0], "exponents": [-1.0],
                                                   "coefficients": [[1.0]]}]}}}
    with pytest.raises(ValueError, match="Invalid basis shell for H"):
        Molecule([("H", (0, 0, 0))], sto)
