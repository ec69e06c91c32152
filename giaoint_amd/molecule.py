"""Molecules for the ERI and Cholesky entry points.

Mirrors the part of python/giao_integrals/molecule.py that feeds the ERI
path: geometry in angstrom or bohr (1 bohr = 0.529177210544 angstrom,
molecule.py:26,65-83), XYZ input (molecule.py:86-112), the basis built from
Basis Set Exchange JSON in atom order and then stable increasing L order
within each atom, a general contraction kept as one shell with several
coefficient rows and combined (e.g. SP) shells split per angular momentum
(molecule.py:191-267), and the eri / cholesky_eri / eri_batches methods
(molecule.py:390-450). Named bases and Gaussian94/NWChem text need the
basis-set-exchange package, which is not a dependency here: pass BSE JSON
(a dictionary, a .json path or JSON text). One-electron integrals are not
part of this package.
"""

from __future__ import annotations

import json
import math
from os import PathLike
from pathlib import Path

from . import cholesky as _cholesky
from . import eri as _eri
from .core import Basis, MagneticField, Shell

BOHR_IN_ANGSTROM = 0.529177210544

_ELEMENTS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
    "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf "
    "Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og"
).split()
_Z = {symbol: z for z, symbol in enumerate(_ELEMENTS, start=1)}


_UNIT_SCALE = {"angstrom": 1.0 / BOHR_IN_ANGSTROM, "bohr": 1.0}
_FORTRAN_EXPONENT = str.maketrans("Dd", "Ee")


def _real(value) -> float:
    """A float from a number or from numeric text (Fortran D exponents allowed)."""
    if isinstance(value, bool):
        raise TypeError("expected a number")
    if isinstance(value, str):
        value = value.translate(_FORTRAN_EXPONENT)
    return float(value)


def _element(value) -> str:
    """Element symbol from a symbol (any case) or an atomic number."""
    if isinstance(value, str):
        text = value.strip()
        if text.isdigit():
            z = int(text)
        else:
            symbol = text[:1].upper() + text[1:].lower()
            if symbol in _Z:
                return symbol
            z = -1
    elif isinstance(value, int) and not isinstance(value, bool):
        z = value
    else:
        z = -1
    if 1 <= z <= len(_ELEMENTS):
        return _ELEMENTS[z - 1]
    raise ValueError(f"Unknown element {value!r}; use an element symbol or atomic number")


def _atom_record(entry, scale: float):
    """(symbol, position in bohr) of one (symbol, (x, y, z)) entry; the
    coordinates are checked before the symbol (molecule.py:65-83)."""
    symbol, coordinates = entry
    position = tuple(scale * _real(c) for c in coordinates)
    if len(position) != 3 or not all(map(math.isfinite, position)):
        raise ValueError("coordinates must be three finite numbers")
    return _element(symbol), position


def _atoms_in_bohr(atoms, unit: str):
    scale = _UNIT_SCALE.get(unit) if isinstance(unit, str) else None
    if scale is None:
        raise ValueError("unit must be 'angstrom' or 'bohr'")
    records = []
    for number, entry in enumerate(atoms, start=1):
        try:
            records.append(_atom_record(entry, scale))
        except (TypeError, ValueError, OverflowError) as error:
            raise ValueError(f"Invalid atom {number}: {error}") from error
    if not records:
        raise ValueError("A molecule must contain at least one atom")
    return tuple(records)


def _xyz_text(source) -> str:
    """Text of an XYZ source: multiline text as given (a leading BOM removed),
    anything else a path read as UTF-8 (with or without BOM)."""
    if isinstance(source, str) and ("\n" in source or "\r" in source):
        return source.removeprefix("\ufeff")
    return Path(source).read_text(encoding="utf-8-sig")


def _xyz_atoms(source):
    """Atoms of one XYZ frame (molecule.py:86-112): a count line, a comment
    line, the atom lines, and nothing but blank lines after them. Blank lines
    before the count are ignored; line numbers in messages count from it."""
    rows = _xyz_text(source).splitlines()
    first = next((k for k, row in enumerate(rows) if row.strip()), len(rows))
    header, body = rows[first:first + 1], rows[first + 1:]
    try:
        count = int(header[0])
    except (IndexError, ValueError) as error:
        raise ValueError("XYZ must start with an integer atom count") from error
    if count < 1:
        raise ValueError("XYZ atom count must be positive")
    frame, rest = body[1:count + 1], body[count + 1:]
    if len(frame) < count:
        raise ValueError(f"XYZ requires a comment line and {count} atom lines")
    if any(row.strip() for row in rest):
        raise ValueError("Expected one XYZ frame; extra atoms or frames follow the declared count")
    atoms = []
    for line_number, row in zip(range(3, count + 3), frame):
        fields = row.split()
        if len(fields) != 4:
            raise ValueError(f"XYZ line {line_number} must contain an element and three coordinates")
        atoms.append((fields[0], fields[1:]))
    return atoms


def _basis_json(source, basis_format):
    """BSE JSON data from a dictionary, a path or JSON text."""
    if basis_format not in (None, "json", "gaussian94", "nwchem"):
        raise ValueError("basis_format must be 'json', 'gaussian94', or 'nwchem'")
    if isinstance(source, dict):
        if basis_format not in (None, "json"):
            raise ValueError("A basis dictionary uses basis_format='json'")
        return source
    if not isinstance(source, (str, PathLike)):
        raise TypeError("basis must be a basis name, file path, formatted text, or BSE dictionary")
    if basis_format in ("gaussian94", "nwchem"):
        raise ImportError("Gaussian94 and NWChem bases require basis-set-exchange, which is not available; "
                          "pass BSE JSON instead")
    if isinstance(source, str) and source.lstrip().startswith("{"):
        text = source
    else:
        path = Path(source)
        if not (isinstance(source, PathLike) or path.is_file() or path.suffix or "/" in str(source)):
            raise ImportError(f"Named bases ({source!r}) require basis-set-exchange, which is not available; "
                              "pass BSE JSON instead")
        if path.suffix.lower() not in ("", ".json") and basis_format is None:
            raise ImportError("Gaussian94 and NWChem bases require basis-set-exchange, which is not available; "
                              "pass BSE JSON instead")
        text = path.read_text(encoding="utf-8-sig")
    try:
        data = json.loads(text)
    except ValueError as error:
        raise ValueError(f"Invalid BSE JSON basis: {error}") from error
    if not isinstance(data, dict):
        raise ValueError("BSE JSON basis must be an object containing 'elements'")
    return data


_SHELL_TYPES = frozenset({"gto", "gto_cartesian", "gto_spherical"})


def _entry_shells(entry):
    """(L, exponents, coefficient rows) of one BSE shell entry; a combined
    shell (e.g. SP) gives one shell per angular momentum, each with its own
    coefficient row, and a general contraction stays one shell."""
    if entry.get("function_type", "gto") not in _SHELL_TYPES:
        raise ValueError("only Gaussian orbital shells are supported")
    momenta = entry["angular_momentum"]
    if not momenta or any(type(l) is not int or l < 0 for l in momenta):
        raise ValueError("angular momenta must be non-negative integers")
    if len(set(momenta)) != len(momenta):
        raise ValueError("combined shells must have distinct angular momenta")
    exponents = [_real(x) for x in entry["exponents"]]
    rows = [[_real(x) for x in row] for row in entry["coefficients"]]
    if len(momenta) == 1:
        return [(momenta[0], exponents, rows)]
    if len(rows) != len(momenta):
        raise ValueError("combined shells require one coefficient row per angular momentum")
    return [(l, exponents, [row]) for l, row in zip(momenta, rows)]


def _element_shells(elements, symbol):
    """Shell specifications of one element in increasing L, stable within
    equal L (the order of the basis file), molecule.py:191-267."""
    element = elements.get(str(_Z[symbol]))
    if not isinstance(element, dict):
        raise ValueError(f"Basis has no data for element {symbol}")
    if element.get("ecp_potentials") or element.get("ecp_electrons", 0):
        raise ValueError(f"Basis for {symbol} contains an ECP; only all-electron bases are supported")
    entries = element.get("electron_shells")
    if not isinstance(entries, list) or not entries:
        raise ValueError(f"Basis for {symbol} has no electron shells")
    specs = []
    for index, entry in enumerate(entries):
        try:
            specs.extend(_entry_shells(entry))
        except (AttributeError, KeyError, TypeError, ValueError, OverflowError) as error:
            raise ValueError(f"Invalid basis shell {index} for {symbol}: {error}") from error
    specs.sort(key=lambda spec: spec[0])
    return specs


def _shells(atoms, data):
    """The basis in atom order (each element's shells parsed once) and the
    owning atom of every shell."""
    elements = data.get("elements")
    if not isinstance(elements, dict):
        raise ValueError("BSE basis data must contain an 'elements' object keyed by atomic number")
    parsed = {}
    shells, owners = [], []
    for owner, (symbol, position) in enumerate(atoms):
        if symbol not in parsed:
            parsed[symbol] = _element_shells(elements, symbol)
        for l, exponents, rows in parsed[symbol]:
            try:
                shells.append(Shell(position, l, exponents, rows))
            except (TypeError, ValueError, OverflowError) as error:
                raise ValueError(f"Invalid basis shell for {symbol}: {error}") from error
            owners.append(owner)
    return Basis(shells), tuple(owners)


class Molecule:
    """Geometry, a Cartesian basis and the ERI methods (molecule.py:285-450).

    ``atoms`` is ``[(symbol, (x, y, z)), ...]`` in angstrom unless
    ``unit='bohr'``; stored coordinates are in bohr. ``basis`` is BSE JSON
    (dictionary, path or text). Output is Cartesian.
    """

    def __init__(self, atoms, basis, *, unit: str = "angstrom", basis_format=None, field=None):
        if field is not None and not isinstance(field, MagneticField):
            raise TypeError("field must be a MagneticField or None")
        self._atoms = _atoms_in_bohr(atoms, unit)
        self._basis, self._shell_atoms = _shells(self._atoms, _basis_json(basis, basis_format))
        self._field = field if field is not None else MagneticField()

    @classmethod
    def from_xyz(cls, source, basis, *, unit: str = "angstrom", basis_format=None, field=None) -> "Molecule":
        """One XYZ frame from a path or multiline text."""
        return cls(_xyz_atoms(source), basis, unit=unit, basis_format=basis_format, field=field)

    @property
    def atoms(self):
        return self._atoms

    @property
    def basis(self) -> Basis:
        return self._basis

    @property
    def shell_atoms(self):
        return self._shell_atoms

    @property
    def field(self) -> MagneticField:
        return self._field

    @property
    def ao_count(self) -> int:
        return self._basis.ao_count

    def __repr__(self) -> str:
        return f"Molecule(atoms={len(self._atoms)}, shells={len(self._shell_atoms)}, ao_count={self.ao_count})"

    def eri(self, *, max_bytes=None, screening_threshold: float = 0.0, threads=None):
        """The full (nao, nao, nao, nao) tensor; max_bytes=None means no cap."""
        cap = self.ao_count**4 * 16 if max_bytes is None else max_bytes
        return _eri.eri(self._basis, field=self._field, storage="full", max_bytes=cap,
                        screening_threshold=screening_threshold, threads=threads)

    def cholesky_eri(self, *, threshold: float = 1e-8, screening_threshold: float = 0.0, threads=None,
                     max_bytes=None, pivot_batch_size: int = 32):
        """Coulomb Cholesky factors (see giaoint_amd.cholesky.cholesky_eri)."""
        return _cholesky.cholesky_eri(self._basis, field=self._field, threshold=threshold,
                                      screening_threshold=screening_threshold, threads=threads,
                                      max_bytes=max_bytes, pivot_batch_size=pivot_batch_size)

    def eri_batches(self, *, target_bytes: int = 64 * 1024**2, screening_threshold: float = 0.0, threads=None):
        """Canonical shell-quartet batches with the molecule's field."""
        return _eri.eri_batches(self._basis, field=self._field, target_bytes=target_bytes,
                                screening_threshold=screening_threshold, threads=threads)

    def eri_derivative_batches(self, *, quartets=None, target_bytes: int = 64 * 1024**2,
                               screening_threshold: float = 0.0, threads=None):
        """ERI centre-derivative batches with the molecule's field; blocks
        (4, 3, na, nb, nc, nd), slots in quartet order (shell_atoms maps
        shells to atoms; repeated slots add)."""
        return _eri.eri_derivative_batches(self._basis, field=self._field, quartets=quartets,
                                           target_bytes=target_bytes, screening_threshold=screening_threshold,
                                           threads=threads)
