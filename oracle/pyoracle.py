"""ctypes access to the CPU checkers (test infrastructure only).

``load("port")`` opens oracle/liboracle.so (the C restatement in
giao_oracle.c); ``load("reference")`` opens oracle/_ref/libgiao_ref.so, which
oracle/Makefile compiles from GIAOINT's own sources. Both export the same
functions (prefix ``or_`` / ``ref_``), so a test can run either as the
checker. Only tests, ``__graft_entry__.smoke`` and bench.py's cpu_baseline
leg import this module; the product package never does.

A basis is given as a list of ``(center, L, exponents, coefficients)`` with
``coefficients`` shaped (ncontr, nprim) and already contraction-normalized
(use :meth:`Checker.normalize_shell`), matching the C layout in giao_oracle.h.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBS = {"port": HERE / "liboracle.so", "reference": HERE / "_ref" / "libgiao_ref.so",
        "reference_fma": HERE / "_ref" / "libgiao_ref_fma.so"}
PREFIX = {"port": "or_", "reference": "ref_", "reference_fma": "ref_"}

ERRORS = {1: ValueError, 2: ValueError, 3: IndexError, 4: ArithmeticError, 5: RuntimeError}
ERROR_NAMES = {1: "invalid_argument", 2: "length_error", 3: "out_of_range", 4: "BoysNumericalError", 5: "runtime"}


class CheckerError(Exception):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{ERROR_NAMES.get(code, code)}] {message}")
        self.code = code
        self.message = message


class _Basis(C.Structure):
    _fields_ = [
        ("nshell", C.c_int),
        ("centers", C.POINTER(C.c_double)),
        ("L", C.POINTER(C.c_int)),
        ("nprim", C.POINTER(C.c_int)),
        ("ncontr", C.POINTER(C.c_int)),
        ("exps", C.POINTER(C.c_double)),
        ("coefs", C.POINTER(C.c_double)),
    ]


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


class BasisArrays:
    """Flat arrays for the C basis layout; keeps the numpy buffers alive."""

    def __init__(self, shells):
        self.shells = [
            (np.asarray(c, float), int(l), np.asarray(e, float), np.atleast_2d(np.asarray(k, float)))
            for c, l, e, k in shells
        ]
        self.centers = np.ascontiguousarray([s[0] for s in self.shells], dtype=np.float64).reshape(-1)
        self.L = np.ascontiguousarray([s[1] for s in self.shells], dtype=np.int32)
        self.nprim = np.ascontiguousarray([len(s[2]) for s in self.shells], dtype=np.int32)
        self.ncontr = np.ascontiguousarray([s[3].shape[0] for s in self.shells], dtype=np.int32)
        self.exps = np.ascontiguousarray(np.concatenate([s[2] for s in self.shells]), dtype=np.float64)
        self.coefs = np.ascontiguousarray(np.concatenate([s[3].reshape(-1) for s in self.shells]), dtype=np.float64)
        self.ao_counts = np.array([ncart(s[1]) * s[3].shape[0] for s in self.shells], dtype=np.int64)
        self.ao_offsets = np.concatenate([[0], np.cumsum(self.ao_counts)[:-1]]).astype(np.int64)
        self.nao = int(self.ao_counts.sum())
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        self.struct = _Basis(
            len(self.shells),
            self.centers.ctypes.data_as(dp),
            self.L.ctypes.data_as(ip),
            self.nprim.ctypes.data_as(ip),
            self.ncontr.ctypes.data_as(ip),
            self.exps.ctypes.data_as(dp),
            self.coefs.ctypes.data_as(dp),
        )

    def quartet_size(self, q) -> int:
        return int(np.prod([self.ao_counts[i] for i in q]))


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class Checker:
    def __init__(self, kind: str = "port"):
        path = LIBS[kind]
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (make -C oracle {'ref' if kind == 'reference' else 'oracle'})")
        self.kind = kind
        self.lib = C.CDLL(str(path))
        p = PREFIX[kind]
        self._err = getattr(self.lib, p + "last_error")
        self._err.restype = C.c_char_p
        self._p = p

    def _fn(self, name):
        f = getattr(self.lib, self._p + name)
        f.restype = C.c_int
        return f

    def _check(self, status):
        if status:
            raise CheckerError(status, self._err().decode())

    @staticmethod
    def field6(B=(0.0, 0.0, 0.0), origin=(0.0, 0.0, 0.0)):
        return np.ascontiguousarray(list(B) + list(origin), dtype=np.float64)

    def normalize_shell(self, L, exps, coefs):
        exps = np.ascontiguousarray(exps, dtype=np.float64)
        coefs = np.ascontiguousarray(np.atleast_2d(coefs), dtype=np.float64)
        out = np.empty_like(coefs)
        self._check(
            self._fn("normalize_shell")(
                C.c_int(L), C.c_int(len(exps)), C.c_int(coefs.shape[0]), _ptr(exps), _ptr(coefs), _ptr(out)
            )
        )
        return out

    def boys(self, z, nmax, scaled=False):
        out = np.empty(2 * (nmax + 1))
        region = C.c_int()
        err = C.c_double()
        seg = C.c_int()
        self._check(
            self._fn("boys")(
                C.c_double(complex(z).real),
                C.c_double(complex(z).imag),
                C.c_int(nmax),
                C.c_int(int(scaled)),
                _ptr(out),
                C.byref(region),
                C.byref(err),
                C.byref(seg),
            )
        )
        return out.view(np.complex128), {"region": region.value, "error": err.value, "segments": seg.value}

    def primitive_eri(self, prims, field):
        """prims: four dicts with exponent, center, angular, coefficient, normalized."""
        p = np.zeros((4, 6))
        a = np.zeros((4, 3), dtype=np.int32)
        for i, d in enumerate(prims):
            p[i] = [d["exponent"], *d["center"], d.get("coefficient", 1.0), float(d.get("normalized", True))]
            a[i] = d.get("angular", (0, 0, 0))
        out = np.zeros(2)
        f = np.ascontiguousarray(field, dtype=np.float64)
        self._check(self._fn("primitive_eri")(_ptr(p), _ptr(a, C.c_int), _ptr(f), _ptr(out)))
        return complex(out[0], out[1])

    def eri_batch(self, basis: BasisArrays, quartets, field, threads=1):
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.uint32).reshape(-1, 4))
        sizes = np.array([basis.quartet_size(x) for x in q], dtype=np.uint64)
        offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
        out = np.zeros(2 * int(offsets[-1]) if len(q) else 2)
        f = np.ascontiguousarray(field, dtype=np.float64)
        self._check(
            self._fn("eri_batch")(
                C.byref(basis.struct),
                _ptr(q, C.c_uint32),
                C.c_size_t(len(q)),
                _ptr(f),
                _ptr(offsets, C.c_uint64),
                _ptr(out),
                C.c_int(threads),
            )
        )
        return out[: 2 * int(offsets[-1])].view(np.complex128), offsets

    def eri_derivative_blocks(self, basis: BasisArrays, quartets, field, threads=1, offsets=None, out_len=None):
        """compute_eri_derivative_blocks: packed (4, 3, na, nb, nc, nd) blocks
        and the n + 1 offsets (complex elements)."""
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.uint32).reshape(-1, 4))
        if offsets is None:
            sizes = np.array([12 * basis.quartet_size(x) for x in q], dtype=np.uint64)
            offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        n_out = int(offsets[-1]) if out_len is None else int(out_len)
        out = np.zeros(2 * max(n_out, 1))
        f = np.ascontiguousarray(field, dtype=np.float64)
        self._check(
            self._fn("eri_derivative_blocks")(
                C.byref(basis.struct), _ptr(q, C.c_uint32), C.c_size_t(len(q)), _ptr(f), _ptr(offsets, C.c_uint64),
                _ptr(out), C.c_size_t(n_out), C.c_int(threads)))
        return out[: 2 * n_out].view(np.complex128), offsets

    def eri_derivative_pair_bounds(self, basis: BasisArrays, field):
        ns = len(basis.L)
        out = np.zeros((ns, ns))
        f = np.ascontiguousarray(field, dtype=np.float64)
        self._check(self._fn("eri_derivative_pair_bounds")(C.byref(basis.struct), _ptr(f), _ptr(out)))
        return out

    def schwarz_bounds(self, basis: BasisArrays, field):
        ns = len(basis.L)
        out = np.zeros((ns, ns))
        f = np.ascontiguousarray(field, dtype=np.float64)
        self._check(self._fn("schwarz_bounds")(C.byref(basis.struct), _ptr(f), _ptr(out)))
        return out

    def eri_tensor(self, basis: BasisArrays, field, threads=1):
        n = basis.nao
        out = np.zeros(2 * n**4)
        f = np.ascontiguousarray(field, dtype=np.float64)
        self._check(self._fn("eri_tensor")(C.byref(basis.struct), _ptr(f), _ptr(out), C.c_int(threads)))
        return out.view(np.complex128).reshape(n, n, n, n)

    def nuclear_matrix(self, basis: BasisArrays, charges, centers, field):
        n = basis.nao
        z = np.ascontiguousarray(charges, dtype=np.float64).reshape(-1)
        c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(2 * n * n)
        f = np.ascontiguousarray(field, dtype=np.float64)
        self._check(self._fn("nuclear_matrix")(C.byref(basis.struct), C.c_int(len(z)), _ptr(z), _ptr(c), _ptr(f),
                                               _ptr(out)))
        return out.view(np.complex128).reshape(n, n)

    def cholesky_diagonal(self, basis: BasisArrays, pairs, field, threads=1):
        n = basis.nao
        p = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint64).reshape(-1, 2))
        out = np.zeros(n * n)
        f = np.ascontiguousarray(field, dtype=np.float64)
        self._check(
            self._fn("cholesky_diagonal")(
                C.byref(basis.struct), _ptr(p, C.c_uint64), C.c_size_t(len(p)), _ptr(f), _ptr(out), C.c_int(threads)
            )
        )
        return out.reshape(n, n)

    def cholesky_columns(self, basis: BasisArrays, pivots, field, threads=1, real=False):
        n = basis.nao
        p = np.ascontiguousarray(np.asarray(pivots, dtype=np.uint64).reshape(-1, 2))
        out = np.zeros(len(p) * n * n * (1 if real else 2))
        f = np.ascontiguousarray(field, dtype=np.float64)
        self._check(
            self._fn("cholesky_columns")(
                C.byref(basis.struct),
                _ptr(p, C.c_uint64),
                C.c_size_t(len(p)),
                _ptr(f),
                _ptr(out),
                C.c_int(int(real)),
                C.c_int(threads),
            )
        )
        return out.reshape(len(p), n * n) if real else out.view(np.complex128).reshape(len(p), n * n)


_CACHE: dict[str, Checker] = {}


def load(kind: str = "port") -> Checker:
    if kind not in _CACHE:
        _CACHE[kind] = Checker(kind)
    return _CACHE[kind]


def available(kind: str) -> bool:
    return LIBS[kind].exists()


def make_basis(shells, checker: Checker | None = None, normalize=True) -> BasisArrays:
    """Build BasisArrays from (center, L, exps, raw coefficients) tuples."""
    chk = checker or load("port")
    out = []
    for c, l, e, k in shells:
        k = np.atleast_2d(np.asarray(k, float))
        out.append((c, l, e, chk.normalize_shell(l, e, k) if normalize else k))
    return BasisArrays(out)
