"""The ERI entry points of GIAOINT's extension module, over the HIP library.

Same names, arguments, results and exceptions as the reference's pybind11
module for the electron-repulsion path (src/bindings.cpp:293-485,599-764,
2255-2330): MagneticField, PrimitiveGaussian, Shell, Basis, BoysRegion,
boys, boys_with_diagnostics, primitive_eri, eri_shell, _eri_full, _eri_batch,
_EriSchwarzBounds, _canonical_shell_quartet, _cholesky_diagonal,
_cholesky_columns, _build_info, openmp_enabled and openmp_max_threads.

Every integral is evaluated on the GPU through libgiaohip.so (one context on
device 0, created on first use). ``threads`` is validated as the reference
validates it and otherwise has no effect: results never depend on it.
"""

from __future__ import annotations

import enum
import operator
import threading
import weakref

import numpy as np

from . import _native

__all__ = [
    "MagneticField",
    "PrimitiveGaussian",
    "Shell",
    "Basis",
    "BoysRegion",
    "boys",
    "boys_with_diagnostics",
    "primitive_eri",
    "eri_shell",
    "_eri_full",
    "_eri_batch",
    "_EriSchwarzBounds",
    "eri_center_derivatives_shell",
    "_eri_derivative_batch",
    "_EriDerivativeSchwarzBounds",
    "_canonical_shell_quartet",
    "_cholesky_diagonal",
    "_cholesky_columns",
    "_build_info",
    "openmp_enabled",
    "openmp_max_threads",
    "BoysNumericalError",
]

BoysNumericalError = _native.BoysNumericalError

_ctx = None
_ctx_lock = threading.Lock()


def _context() -> _native.Context:
    global _ctx
    with _ctx_lock:
        if _ctx is None:
            _ctx = _native.Context(0)
        return _ctx


def _vec3(value, name: str) -> tuple:
    a = np.asarray(value)
    if a.ndim != 1 or a.shape[0] != 3:
        raise ValueError(f"{name} must be a one-dimensional array with length 3")
    return tuple(float(x) for x in a.astype(np.float64))


def _threads(value) -> int:
    """_threads.py:21-30 validation; the GPU path ignores the count."""
    if value is None:
        return 1
    if isinstance(value, (bool, np.bool_)):
        raise ValueError("threads must be a positive integer")
    try:
        n = operator.index(value)
    except TypeError as error:
        raise ValueError("threads must be a positive integer") from error
    if not 1 <= n <= 2**31 - 1:
        raise ValueError("threads must be a positive integer <= 2147483647")
    return n


class MagneticField:
    """Uniform field B and gauge origin (types.hpp:126-132)."""

    def __init__(self, B=(0.0, 0.0, 0.0), gauge_origin=(0.0, 0.0, 0.0)):
        self._B = _vec3(B, "B")
        self._O = _vec3(gauge_origin, "gauge_origin")

    @property
    def B(self) -> tuple:
        return self._B

    @property
    def gauge_origin(self) -> tuple:
        return self._O

    def london_wave_vector(self, center) -> tuple:
        """kappa = 1/2 B x (C - O) (types.cpp:179-186)."""
        c = _vec3(center, "center")
        d = [c[i] - self._O[i] for i in range(3)]
        B = self._B
        return (0.5 * (B[1] * d[2] - B[2] * d[1]), 0.5 * (B[2] * d[0] - B[0] * d[2]),
                0.5 * (B[0] * d[1] - B[1] * d[0]))

    def _field6(self) -> np.ndarray:
        return _native.field6(self._B, self._O)


def _field6(field) -> np.ndarray:
    return _native.field6() if field is None else field._field6()


def _angular(value) -> tuple:
    if isinstance(value, (str, bytes)) or not hasattr(value, "__len__") or len(value) != 3:
        raise ValueError("angular momentum must contain exactly three integers")
    out = []
    for v in value:
        if isinstance(v, (bool, np.bool_)):
            raise ValueError("angular momentum must contain exactly three integers")
        try:
            k = operator.index(v)
        except TypeError as error:
            raise ValueError("angular momentum must contain exactly three integers") from error
        if k < 0:
            raise ValueError("angular momentum components must be non-negative")
        out.append(k)
    return tuple(out)


class PrimitiveGaussian:
    """One Cartesian primitive (types.hpp:40-55)."""

    def __init__(self, exponent, center, angular=(0, 0, 0), coefficient=1.0, normalized=True):
        self.exponent = float(exponent)
        if not np.isfinite(self.exponent) or self.exponent <= 0.0:
            raise ValueError("Gaussian exponent must be finite and positive")
        self.center = _vec3(center, "center")
        self.angular = _angular(angular)
        self.coefficient = float(coefficient)
        if not np.isfinite(self.coefficient):
            raise ValueError("Gaussian coefficient must be finite")
        self.normalized = bool(normalized)

    def _row(self) -> list:
        return [self.exponent, *self.center, self.coefficient, float(self.normalized)]


def _cartesian(l: int) -> tuple:
    return tuple((x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1))


class Shell:
    """A generally contracted Cartesian shell (types.hpp:57-106).

    ``coefficients`` is [contraction][primitive] (or one contraction as a
    one-dimensional array). With ``normalize`` (the default) the
    contractions are normalized as in types.cpp:239-306.
    """

    def __init__(self, center, angular_momentum, exponents, coefficients, *, normalize=True):
        self._center = _vec3(center, "center")
        try:
            l = operator.index(angular_momentum)
        except TypeError as error:
            raise ValueError("angular momentum must be a non-negative integer") from error
        if l < 0:
            raise ValueError("angular momentum must be a non-negative integer")
        e = np.asarray(exponents, dtype=np.float64)
        if e.ndim != 1:
            raise ValueError("exponents must be a one-dimensional array")
        c = np.asarray(coefficients, dtype=np.float64)
        if c.ndim not in (1, 2):
            raise ValueError("coefficients must be one- or two-dimensional")
        c = c.reshape(1, -1) if c.ndim == 1 else c
        if c.shape[1] != e.size:
            raise ValueError("the final coefficient dimension must equal the exponent count")
        self._l = int(l)
        self._exps = np.ascontiguousarray(e)
        raw = np.ascontiguousarray(c)
        # validation (and normalization) by the library's host code
        normalized = _native.normalize_shells([self._l], [e.size], [raw.shape[0]], self._exps, raw.reshape(-1))
        self._coefs = normalized.reshape(raw.shape) if normalize else raw.copy()

    @property
    def center(self) -> tuple:
        return self._center

    @property
    def angular_momentum(self) -> int:
        return self._l

    @property
    def primitive_count(self) -> int:
        return int(self._exps.size)

    @property
    def contraction_count(self) -> int:
        return int(self._coefs.shape[0])

    @property
    def cartesian_count(self) -> int:
        return (self._l + 1) * (self._l + 2) // 2

    @property
    def ao_count(self) -> int:
        return self.cartesian_count * self.contraction_count

    @property
    def components(self) -> tuple:
        return _cartesian(self._l)


class Basis:
    """An ordered list of shells (types.hpp:108-124); uploaded to the GPU on first use."""

    def __init__(self, shells):
        self._shells = tuple(shells)
        for s in self._shells:
            if not isinstance(s, Shell):
                raise TypeError("Basis needs Shell objects")
        counts = [s.ao_count for s in self._shells]
        self._offsets = tuple(int(x) for x in np.concatenate([[0], np.cumsum(counts)])[:-1]) if counts else ()
        self._nao = int(sum(counts))
        self._device = None
        self._lock = threading.Lock()

    @property
    def shells(self) -> list:
        return list(self._shells)

    @property
    def ao_offsets(self) -> tuple:
        return self._offsets

    @property
    def ao_count(self) -> int:
        return self._nao

    def _dev(self) -> _native.DeviceBasis:
        with self._lock:
            if self._device is None:
                sh = self._shells
                self._device = _native.DeviceBasis(
                    _context(),
                    np.array([s.center for s in sh], dtype=np.float64).reshape(-1),
                    [s.angular_momentum for s in sh],
                    [s.primitive_count for s in sh],
                    [s.contraction_count for s in sh],
                    np.concatenate([s._exps for s in sh]) if sh else np.zeros(0),
                    np.concatenate([s._coefs.reshape(-1) for s in sh]) if sh else np.zeros(0),
                )
            return self._device


class BoysRegion(enum.IntEnum):
    """boys.hpp:17-22"""

    POWER_SERIES = 0
    ADAPTIVE_QUADRATURE = 1
    SCALED_QUADRATURE = 2
    POSITIVE_ASYMPTOTIC = 3


def _order(maximum_order) -> int:
    if isinstance(maximum_order, (bool, np.bool_)):
        raise TypeError("maximum_order must be an integer")
    n = operator.index(maximum_order)
    if n < 0:
        raise TypeError("maximum_order must be a non-negative integer")
    return n


def boys_with_diagnostics(argument, maximum_order, *, scaled=False):
    """F_0..F_maximum_order(argument) and the region diagnostics (bindings.cpp:464-485)."""
    n = _order(maximum_order)
    vals, region, err, seg = _context().boys([complex(argument)], n, scaled=bool(scaled))
    details = {
        "region": BoysRegion(int(region[0])),
        "estimated_absolute_error": float(err[0]),
        "quadrature_segments": int(seg[0]),
        "scaled": bool(scaled),
    }
    return np.ascontiguousarray(vals[0]), details


def boys(argument, maximum_order, *, scaled=False):
    """F_0..F_maximum_order(argument) (bindings.cpp:444-462; boys.cpp:245-294)."""
    return boys_with_diagnostics(argument, maximum_order, scaled=scaled)[0]


def primitive_eri(a, b, c, d, *, field=None) -> complex:
    """(ab|cd) of four primitives (eri.cpp:438-452)."""
    prims = [{"exponent": p.exponent, "center": p.center, "angular": p.angular, "coefficient": p.coefficient,
              "normalized": p.normalized} for p in (a, b, c, d)]
    return _context().primitive_eri(prims, _field6(field))


def _check_out(out, shape):
    if out is None:
        return np.empty(shape, dtype=np.complex128)
    if not isinstance(out, np.ndarray):
        raise ValueError("out must be a NumPy array")
    if out.dtype != np.complex128:
        raise ValueError("out must have dtype complex128")
    if out.ndim != len(shape):
        raise ValueError("out has an incorrect number of dimensions")
    if tuple(out.shape) != tuple(shape):
        raise ValueError("out has an incorrect shape")
    if not out.flags.c_contiguous:
        raise ValueError("out must be C-contiguous")
    if not out.flags.writeable:
        raise ValueError("out must be writable")
    return out


def eri_shell(a, b, c, d, *, field=None, out=None):
    """Contracted block (ab|cd), shape (a.ao_count, b.ao_count, c.ao_count, d.ao_count) (eri.cpp:454-585)."""
    shells = (a, b, c, d)
    out = _check_out(out, tuple(s.ao_count for s in shells))
    basis = Basis(shells)
    basis._dev().eri_batch([(0, 1, 2, 3)], _field6(field), out=out.reshape(-1))
    return out


def eri_center_derivatives_shell(a, b, c, d, *, field=None, out=None):
    """Centre derivatives of (ab|cd) with the London phase response, shape
    (4, 3, na, nb, nc, nd): quartet slot, axis, AOs (bindings.cpp:1308-1323,
    derivatives.cpp:952-959)."""
    shells = (a, b, c, d)
    out = _check_out(out, (4, 3) + tuple(s.ao_count for s in shells))
    Basis(shells)._dev().eri_derivative_batch([(0, 1, 2, 3)], _field6(field), out=out.reshape(-1))
    return out


class _EriDerivativeSchwarzBounds:
    """Derivative-density Schwarz bounds (derivatives.cpp:1307-1387): per
    shell pair the triangle/Schwarz bound of every centre derivative of the
    pair density, London terms and |contraction coefficients| included; a
    quartet's bound pairs it with the ordinary Schwarz factor of the other
    pair. Both tables come from the device."""

    def __init__(self, basis, *, field=None, threads=None):
        _threads(threads)
        self._basis_ref = weakref.ref(basis)
        self._field = (0.0,) * 6 if field is None else tuple(field._field6())
        dev = basis._dev()
        f = _field6(field)
        self._ordinary = dev.schwarz_bounds(f)
        self._values = dev.eri_derivative_pair_bounds(f)
        for v in (self._ordinary, self._values):
            v.setflags(write=False)

    def quartet_bound(self, quartet) -> float:
        q = tuple(int(i) for i in quartet)
        n = self._values.shape[0]
        if len(q) != 4 or any(i < 0 or i >= n for i in q):
            raise IndexError("derivative shell quartet index is out of range")
        return float(self._bounds(np.asarray([q], dtype=np.int64))[0])

    def _bounds(self, q: np.ndarray) -> np.ndarray:
        """derivatives.cpp:1374-1383 for rows of q (validated indices)."""
        D, Q = self._values, self._ordinary
        a, b, c, d = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
        bound = np.maximum(D[a, b] * Q[c, d], Q[a, b] * D[c, d])
        return np.nextafter(bound * (1.0 + 128.0 * np.finfo(np.float64).eps), np.inf)

    def _matches(self, basis, field) -> bool:
        f = (0.0,) * 6 if field is None else tuple(field._field6())
        return self._basis_ref() is basis and f == self._field


def _eri_derivative_batch(basis, quartets, *, field=None, screening_threshold=0.0, threads=None, bounds=None):
    """Packed centre-derivative blocks (bindings.cpp:2255-2304): returns
    (quartets uint32[m,4], shapes int64[m,6] = (4, 3, na, nb, nc, nd),
    offsets int64[m+1], values complex128[offsets[m]], requested, screened);
    quartets whose derivative bound is below a positive threshold are left out."""
    _threads(threads)
    t = float(screening_threshold)
    if not np.isfinite(t) or t < 0.0:
        raise ValueError("screening_threshold must be finite and non-negative")
    if t > 0.0 and bounds is None:
        raise ValueError("derivative screening requires derivative bounds")
    if bounds is not None and not bounds._matches(basis, field):
        raise ValueError("derivative bounds do not match basis and field")
    q = np.asarray(quartets, dtype=np.int64).reshape(-1, 4)
    if q.size and (q.min() < 0 or q.max() >= len(basis.shells)):
        raise IndexError("derivative shell quartet index is out of range")
    requested = len(q)
    if t > 0.0 and len(q):
        q = q[~(bounds._bounds(q) < t)]
    ao = np.array([s.ao_count for s in basis.shells], dtype=np.int64)
    m = len(q)
    shapes = np.zeros((m, 6), dtype=np.int64)
    if m:
        shapes[:, 0], shapes[:, 1] = 4, 3
        shapes[:, 2:] = ao[q]
    sizes = np.prod(shapes, axis=1)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    kept = np.ascontiguousarray(q, dtype=np.uint32)
    values = basis._dev().eri_derivative_batch(kept, _field6(field)) if m else np.zeros(0, dtype=np.complex128)
    return kept, shapes, offsets, values, requested, requested - m


class _EriSchwarzBounds:
    """Shell-pair Schwarz factors (eri.cpp:587-635)."""

    def __init__(self, basis, *, field=None):
        self._basis_ref = weakref.ref(basis)
        self._field = (0.0,) * 6 if field is None else tuple(field._field6())
        self._values = basis._dev().schwarz_bounds(_field6(field))
        self._values.setflags(write=False)

    def bound(self, a, b) -> float:
        n = self._values.shape[0]
        if not (0 <= a < n and 0 <= b < n):
            raise IndexError("ERI Schwarz shell-pair index is out of range")
        return float(self._values[a, b])

    @property
    def values(self) -> np.ndarray:
        return self._values.copy()

    def _matches(self, basis, field) -> bool:
        f = (0.0,) * 6 if field is None else tuple(field._field6())
        return self._basis_ref() is basis and f == self._field


def _screening(threshold) -> float:
    t = float(threshold)
    if not np.isfinite(t) or t < 0.0:
        raise ValueError("ERI screening threshold must be finite and non-negative")
    return t


def _eri_batch(basis, quartets, *, field=None, screening_threshold=0.0, threads=None, bounds=None):
    """Packed blocks of the computed quartets (bindings.cpp:266-282,672-731):
    returns (quartets uint32[m,4], shapes int64[m,4], offsets int64[m+1],
    values complex128[offsets[m]], requested, screened); block i is
    values[offsets[i]:offsets[i+1]]; screened quartets are left out."""
    _threads(threads)
    t = _screening(screening_threshold)
    q = np.ascontiguousarray(np.asarray(quartets, dtype=np.int64).reshape(-1, 4))
    if q.size and (q.min() < 0 or q.max() >= len(basis.shells)):
        raise IndexError("ERI shell-quartet index is out of range")
    q = q.astype(np.uint32)
    if t > 0.0 and bounds is None:
        raise ValueError("positive ERI screening requires Schwarz bounds")
    if bounds is not None and not bounds._matches(basis, field):
        raise ValueError("ERI Schwarz bounds do not match the basis")
    dev = basis._dev()
    kept = dev.screen(q, None if bounds is None else bounds._values, t)
    ao = np.array([s.ao_count for s in basis.shells], dtype=np.int64)
    shapes = ao[kept.astype(np.int64)] if len(kept) else np.zeros((0, 4), dtype=np.int64)
    sizes = np.prod(shapes, axis=1) if len(kept) else np.zeros(0, dtype=np.int64)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)  # m + 1 entries
    values = dev.eri_batch(kept, _field6(field)) if len(kept) else np.zeros(0, dtype=np.complex128)
    return kept, shapes, offsets, values, int(len(q)), int(len(q) - len(kept))


def _eri_full(basis, *, field=None, screening_threshold=0.0, threads=None):
    """Full (n,n,n,n) tensor from canonical quartets and their images (eri.cpp:809-847)."""
    _threads(threads)
    t = _screening(screening_threshold)
    dev = basis._dev()
    if t == 0.0:
        return dev.eri_tensor(_field6(field))
    n = basis.ao_count
    ns = len(basis.shells)
    bounds = _EriSchwarzBounds(basis, field=field)
    canon = sorted({_canonical_shell_quartet(a, b, c, d)[0] for a in range(ns) for b in range(ns)
                    for c in range(ns) for d in range(ns)})
    kept, shapes, offsets, values, _, _ = _eri_batch(basis, canon, field=field, screening_threshold=t, bounds=bounds)
    out = np.zeros((n, n, n, n), dtype=np.complex128)
    off = basis.ao_offsets
    for (a, b, c, d), shp, o in zip(kept.tolist(), shapes.tolist(), offsets[:-1].tolist()):
        blk = values[o:o + int(np.prod(shp))].reshape(shp)
        sa, sb, sc, sd = (slice(off[x], off[x] + basis.shells[x].ao_count) for x in (a, b, c, d))
        out[sa, sb, sc, sd] = blk
        out[sc, sd, sa, sb] = blk.transpose(2, 3, 0, 1)
        out[sb, sa, sd, sc] = blk.transpose(1, 0, 3, 2).conj()
        out[sd, sc, sb, sa] = blk.transpose(3, 2, 1, 0).conj()
    return out


class Nucleus:
    """Point charge (types.hpp:134-139): charge and centre in bohr.

    Rejects what the reference's constructor rejects, with its messages
    (types.cpp:188-194): a charge that is not finite and positive, and a
    centre with a non-finite coordinate."""

    def __init__(self, charge, center):
        charge = float(charge)
        center = _vec3(center, "center")
        if not np.isfinite(charge) or charge <= 0.0:
            raise ValueError("nuclear charge must be finite and positive")
        if not all(np.isfinite(x) for x in center):
            raise ValueError("nuclear center values must be finite")
        self.charge = charge
        self.center = center


def nuclear_attraction(basis, nuclei, *, field=None, out=None):
    """(n, n) complex128 nuclear-attraction matrix, sum over nuclei of
    -Z <mu| 1/|r - C| |nu> (pybind11 `nuclear_attraction`, bindings.cpp:577-596;
    compute_nuclear_attraction_matrix, nuclear.cpp:282-312)."""
    n = basis.ao_count
    if out is not None:
        _check_out(out, (n, n))
    nuclei = list(nuclei)
    z = np.array([nu.charge for nu in nuclei], dtype=np.float64)
    c = np.array([nu.center for nu in nuclei], dtype=np.float64).reshape(-1, 3)
    values = basis._dev().nuclear_attraction(z, c, _field6(field))
    if out is None:
        return values
    out[...] = values
    return out


def _canonical_shell_quartet(a, b, c, d):
    """Lexicographic minimum of {abcd, cdab, badc*, dcba*} and its conjugate flag (eri.cpp:414-429)."""
    images = [((a, b, c, d), False), ((c, d, a, b), False), ((b, a, d, c), True), ((d, c, b, a), True)]
    best = min(images, key=lambda x: x[0])
    return tuple(int(x) for x in best[0]), best[1]


def _pairs(values, name):
    p = np.asarray(values, dtype=np.int64).reshape(-1, 2)
    if p.size and p.min() < 0:
        raise IndexError(f"Cholesky {name} index is out of range")
    return p


def _cholesky_diagonal(basis, pairs, *, field=None, threads=None):
    """Re (mu nu|nu mu) on the given shell pairs a <= b (cholesky.cpp:285-314)."""
    _threads(threads)
    return basis._dev().cholesky_diagonal(_pairs(pairs, "shell-pair"), _field6(field))


def _cholesky_columns(basis, pivots, *, field=None, threads=None):
    """(mu' nu'|nu_P mu_P) rows for pivots sharing one shell pair (cholesky.cpp:203-282);
    float64 rows at exactly zero field, complex128 otherwise."""
    _threads(threads)
    f = _field6(field)
    real = bool(np.all(f[:3] == 0.0))
    return basis._dev().cholesky_columns(_pairs(pivots, "pivot"), f, real=real)


def _build_info() -> dict:
    return {
        "version": "giaoint_amd (gfx950)",
        "compiler": "hipcc",
        "openmp_enabled": True,
        "openmp_runtime_path": None,
        "device": "AMD Instinct MI355X (gfx950)",
    }


def openmp_enabled() -> bool:
    """Parallel evaluation is available (on the GPU), so threads > 1 is accepted."""
    return True


def openmp_max_threads() -> int:
    """Host threads the library plans with (eri.cpp:863-869 reports OpenMP's)."""
    return int(_native.load_library().giaohip_host_threads())
