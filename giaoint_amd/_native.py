"""ctypes binding of the gfx950 C ABI (include/giaohip.h).

The library ``giaoint_amd/libgiaohip.so`` is the only compute path: when it
is missing or no GPU is visible, every call raises. There is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import numpy as np

LIBRARY = Path(__file__).resolve().parent / "libgiaohip.so"

(OK, INVALID_ARGUMENT, LENGTH_ERROR, OUT_OF_RANGE, BOYS_ERROR, DEVICE_ERROR, UNSUPPORTED, OUT_OF_MEMORY,
 RUNTIME_ERROR) = range(9)


class BoysNumericalError(ArithmeticError):
    """Raised where the reference raises giao::BoysNumericalError (boys.hpp:31-34)."""


class GpuUnavailableError(RuntimeError):
    """The HIP library is not built or no GPU is visible."""


_EXC = {
    INVALID_ARGUMENT: ValueError,
    LENGTH_ERROR: ValueError,
    OUT_OF_RANGE: IndexError,
    BOYS_ERROR: BoysNumericalError,
    DEVICE_ERROR: RuntimeError,
    UNSUPPORTED: RuntimeError,
    OUT_OF_MEMORY: MemoryError,
    RUNTIME_ERROR: RuntimeError,
}


class Stats(C.Structure):
    _fields_ = [
        ("requested_quartets", C.c_uint64),
        ("computed_quartets", C.c_uint64),
        ("screened_quartets", C.c_uint64),
        ("primitive_quartets", C.c_uint64),
        ("seed_ms", C.c_double),
        ("transform_ms", C.c_double),
        ("total_ms", C.c_double),
        ("launches", C.c_uint64),
        ("host_ms", C.c_double),
        ("produce_ms", C.c_double),
        ("contract_ms", C.c_double),
        ("contract_launches", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p

# name: (argtypes, restype)
_SIGNATURES = {
    "giaohip_create": ([C.c_int, C.POINTER(_vp)], C.c_int),
    "giaohip_destroy": ([_vp], C.c_int),
    "giaohip_last_error": ([_vp, C.c_char_p, C.c_size_t], C.c_size_t),
    "giaohip_stream": ([_vp], _vp),
    "giaohip_last_stats": ([_vp, C.POINTER(Stats)], C.c_int),
    "giaohip_schwarz_bounds": ([_vp, _vp, _dp, _dp], C.c_int),
    "giaohip_screen_quartets": ([_vp, _vp, _u32p, C.c_size_t, _dp, C.c_double, _u32p, C.POINTER(C.c_size_t)], C.c_int),
    "giaohip_set_option": ([_vp, C.c_int, C.c_int], C.c_int),
    "giaohip_device_alloc": ([_vp, C.c_size_t, C.POINTER(_vp)], C.c_int),
    "giaohip_device_free": ([_vp, _vp], C.c_int),
    "giaohip_memcpy_to_host": ([_vp, _vp, _vp, C.c_size_t], C.c_int),
    "giaohip_synchronize": ([_vp], C.c_int),
    "giaohip_host_threads": ([], C.c_int),
    "giaohip_device_allocations": ([], C.c_ulonglong),
    "giaohip_basis_create": ([_vp, C.c_int, _dp, _ip, _ip, _ip, _dp, _dp, C.POINTER(_vp)], C.c_int),
    "giaohip_basis_update": ([_vp, _vp, C.c_int, _dp, _ip, _ip, _ip, _dp, _dp], C.c_int),
    "giaohip_basis_destroy": ([_vp, _vp], C.c_int),
    "giaohip_basis_ao_count": ([_vp], C.c_size_t),
    "giaohip_normalize_shells": ([C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, C.c_char_p, C.c_size_t], C.c_int),
    "giaohip_primitive_normalization": ([C.c_double, C.c_int, C.c_int, C.c_int], C.c_double),
    "giaohip_boys": ([_vp, _dp, C.c_size_t, C.c_int, C.c_int, _dp, _ip, _dp, _ip], C.c_int),
    "giaohip_eri_batch": ([_vp, _vp, _u32p, C.c_size_t, _dp, _u64p, _vp, C.c_int], C.c_int),
    "giaohip_eri_derivative_batch": ([_vp, _vp, _u32p, C.c_size_t, _dp, _u64p, _vp, C.c_uint64, C.c_int], C.c_int),
    "giaohip_eri_derivative_pair_bounds": ([_vp, _vp, _dp, _dp], C.c_int),
    "giaohip_eri_batch_size": ([_vp, _u32p, C.c_size_t, _u64p], C.c_int),
    "giaohip_eri_batch_noalloc": ([_vp, _vp, _u32p, C.c_size_t, _dp, _u64p, _vp], C.c_int),
    "giaohip_eri_tensor": ([_vp, _vp, _dp, _vp, C.c_int], C.c_int),
    "giaohip_primitive_eri": ([_vp, _dp, _ip, _dp, _dp], C.c_int),
    "giaohip_nuclear_attraction": ([_vp, _vp, C.c_int, _dp, _dp, _dp, _vp], C.c_int),
    "giaohip_cholesky_diagonal": ([_vp, _vp, _u64p, C.c_size_t, _dp, _dp], C.c_int),
    "giaohip_cholesky_columns": ([_vp, _vp, _u64p, C.c_size_t, _dp, _dp, C.c_int], C.c_int),
    "giaohip_cholesky_run": ([_vp, _vp, _dp, C.c_double, C.c_uint64, C.c_uint64, C.POINTER(_vp)], C.c_int),
    "giaohip_cholesky_info": ([_vp, _u64p, _dp, _dp, _ip, _dp, _dp, _dp], C.c_int),
    "giaohip_cholesky_remeasure": ([_vp, _vp, _dp], C.c_int),
    "giaohip_cholesky_pivots": ([_vp, C.POINTER(C.c_int64)], C.c_int),
    "giaohip_cholesky_vectors": ([_vp, _vp, C.c_uint64, C.c_uint64, _dp], C.c_int),
    "giaohip_cholesky_resume": ([_vp, _vp, C.c_uint64], C.c_int),
    "giaohip_cholesky_blocks": ([_vp, _u64p, _u64p, _dp, _u64p, _u64p], C.c_int),
    "giaohip_cholesky_graphs": ([_vp, _u64p, _u64p, _u64p], C.c_int),
    "giaohip_cholesky_reconstruct": ([_vp, _vp, C.POINTER(C.c_int64), _dp], C.c_int),
    "giaohip_cholesky_coulomb": ([_vp, _vp, _dp, _dp], C.c_int),
    "giaohip_cholesky_exchange": ([_vp, _vp, _dp, _dp], C.c_int),
    "giaohip_cholesky_destroy": ([_vp, _vp], C.c_int),
}

EXPORTED = tuple(_SIGNATURES)

_lib = None
_lock = threading.Lock()


def load_library(path: Path | None = None):
    """Load libgiaohip.so and declare its signatures (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path else LIBRARY
        if not p.exists():
            raise GpuUnavailableError(
                f"{p} is not built; run `python -m giaoint_amd.build` (hipcc, gfx950)"
            )
        lib = C.CDLL(str(p))
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        if path is None:
            _lib = lib
        return lib


def normalize_shells(L, nprim, ncontr, exps, coefs) -> np.ndarray:
    """Contraction-normalize many shells at once (host code in the library)."""
    lib = load_library()
    L = np.ascontiguousarray(L, dtype=np.int32)
    nprim = np.ascontiguousarray(nprim, dtype=np.int32)
    ncontr = np.ascontiguousarray(ncontr, dtype=np.int32)
    exps = np.ascontiguousarray(exps, dtype=np.float64)
    coefs = np.ascontiguousarray(coefs, dtype=np.float64)
    out = np.empty_like(coefs)
    msg = C.create_string_buffer(256)
    st = lib.giaohip_normalize_shells(
        int(L.size), _p(L, C.c_int), _p(nprim, C.c_int), _p(ncontr, C.c_int), _p(exps), _p(coefs), _p(out), msg, 256
    )
    if st != OK:
        raise ValueError(msg.value.decode())
    return out


def primitive_normalization(exponent: float, angular) -> float:
    if not np.isfinite(exponent) or exponent <= 0.0:
        raise ValueError("Gaussian exponent must be finite and positive")
    ax, ay, az = (int(v) for v in angular)
    return float(load_library().giaohip_primitive_normalization(float(exponent), ax, ay, az))


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def field6(B=(0.0, 0.0, 0.0), origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    f = np.ascontiguousarray([*map(float, B), *map(float, origin)], dtype=np.float64)
    if f.shape != (6,):
        raise ValueError("field needs three B components and three origin coordinates")
    return f


class Context:
    """One HIP stream on one device (giaohip_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        handle = _vp()
        st = self.lib.giaohip_create(int(device), C.byref(handle))
        if st != OK:
            raise GpuUnavailableError(f"giaohip_create(device={device}) failed with status {st}: no usable GPU")
        self.handle = handle
        self.device = device

    def close(self):
        if getattr(self, "handle", None):
            self.lib.giaohip_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, status: int):
        if status == OK:
            return
        buf = C.create_string_buffer(1024)
        self.lib.giaohip_last_error(self.handle, buf, 1024)
        raise _EXC.get(status, RuntimeError)(buf.value.decode())

    def set_variant(self, variant: str):
        """'batched' (class-specialized kernels where available) or 'plain'."""
        value = {"plain": 0, "batched": 1}[variant]
        self.check(self.lib.giaohip_set_option(self.handle, 0, value))

    def set_chunk_mib(self, mib: int):
        """Work memory of one batched-path chunk in MiB (16..65536)."""
        self.check(self.lib.giaohip_set_option(self.handle, 1, int(mib)))

    def set_cholesky_block(self, candidates: int):
        """Candidates per pivot block of later Cholesky runs (1..32; results do not depend on it)."""
        self.check(self.lib.giaohip_set_option(self.handle, 2, int(candidates)))

    def set_staging_mib(self, mib: int):
        """Device staging of host-memory ERI outputs in MiB (1..65536)."""
        self.check(self.lib.giaohip_set_option(self.handle, 4, int(mib)))

    def set_cholesky_pivot_batch(self, value: int):
        """cholesky_eri's pivot_batch_size for later Cholesky runs (>= 1)."""
        self.check(self.lib.giaohip_set_option(self.handle, 3, int(value)))

    QUEUE_FORMS = {"none": 0, "buckets": 1, "single": 2}

    def set_eri_queue(self, form: str):
        """Launch form of the generic transform kernel: 'none' (one workgroup
        per item and LDS-size bucket), 'buckets' (a persistent device work
        queue per bucket) or 'single' (one persistent launch). Results are
        bitwise identical for every form."""
        self.check(self.lib.giaohip_set_option(self.handle, 5, self.QUEUE_FORMS[form]))

    def set_fast_waves(self, waves: int):
        """Largest workgroup of the class-specialized transform in waves
        (4..8, or 0 for 8; default 4); it sets the bra-pair tile per pass."""
        self.check(self.lib.giaohip_set_option(self.handle, 6, int(waves)))

    MFMA_MODES = {"off": 0, "default": 1, "all": 2, "split32": 3}

    def set_mfma(self, mode: str):
        """Classes on the FP64 matrix cores under the batched variant: "off",
        "default" (classes from the total momentum set by set_mfma_min_l,
        where they are measured faster), "all" (every class with an MFMA instance) or "split32" (as
        "all", with the (gg|gg)- and (ff|dd)-type pair momenta (8, 8) and
        (6, 4) in split FP32 -- a precision study, not within the 1e-11 bar)."""
        self.check(self.lib.giaohip_set_option(self.handle, 7, self.MFMA_MODES[mode]))

    def set_wave_max_l(self, value: int):
        """Largest la + lb + lc + ld whose generic-transform items run on
        one-wavefront workgroups (-1: none). Bitwise identical results."""
        self.check(self.lib.giaohip_set_option(self.handle, 9, int(value)))

    def set_mfma_variant(self, variant: int):
        """Compile-time variant (0..3: GEMM unroll, LDS row padding) of the
        matrix-core contraction kernels; bitwise identical results."""
        self.check(self.lib.giaohip_set_option(self.handle, 12, int(variant)))

    def set_small_max_l(self, value: int):
        """Largest la + lb + lc + ld (-1..2) evaluated by the
        lane-per-primitive-quartet kernel."""
        self.check(self.lib.giaohip_set_option(self.handle, 13, int(value)))

    def set_cholesky_graphs(self, on):
        """Cholesky column launches as hipGraphs: True (every pivot), "inside"
        (pivots inside a block only) or False (launched directly)."""
        self.check(self.lib.giaohip_set_option(self.handle, 11, 2 if on == "inside" else (1 if on else 0)))

    def set_mfma_min_l(self, value: int):
        """Smallest la + lb + lc + ld whose classes (without a class-specialized
        kernel) use the FP64 matrix cores in the "default" MFMA mode."""
        self.check(self.lib.giaohip_set_option(self.handle, 10, int(value)))

    PIPELINES = {"overlap": 0, "serial": 1}

    def set_pipeline(self, mode: str):
        """Producer scheduling of the batched path: "overlap" (second stream)
        or "serial" (context stream)."""
        self.check(self.lib.giaohip_set_option(self.handle, 8, self.PIPELINES[mode]))

    def stats(self) -> dict:
        s = Stats()
        self.check(self.lib.giaohip_last_stats(self.handle, C.byref(s)))
        return s.as_dict()

    # --- device memory -------------------------------------------------
    def alloc(self, nbytes: int) -> int:
        p = _vp()
        self.check(self.lib.giaohip_device_alloc(self.handle, C.c_size_t(int(nbytes)), C.byref(p)))
        return p.value

    def free(self, ptr: int):
        self.check(self.lib.giaohip_device_free(self.handle, _vp(ptr)))

    def to_host(self, ptr: int, out: np.ndarray):
        self.check(self.lib.giaohip_memcpy_to_host(self.handle, out.ctypes.data_as(_vp), _vp(ptr), out.nbytes))
        return out

    def synchronize(self):
        self.check(self.lib.giaohip_synchronize(self.handle))

    # --- Boys ------------------------------------------------------------
    def boys(self, z, nmax: int, scaled: bool = False):
        z = np.ascontiguousarray(np.atleast_1d(np.asarray(z, dtype=np.complex128)))
        n = z.size
        out = np.zeros((n, nmax + 1), dtype=np.complex128) if nmax >= 0 else np.zeros((n, 1), np.complex128)
        region = np.zeros(n, dtype=np.int32)
        err = np.zeros(n)
        seg = np.zeros(n, dtype=np.int32)
        self.check(
            self.lib.giaohip_boys(
                self.handle, _p(z.view(np.float64)), C.c_size_t(n), C.c_int(nmax), C.c_int(int(bool(scaled))),
                _p(out.view(np.float64)), _p(region, C.c_int), _p(err), _p(seg, C.c_int),
            )
        )
        return out, region, err, seg

    def primitive_eri(self, prims, field) -> complex:
        p = np.zeros((4, 6))
        a = np.zeros((4, 3), dtype=np.int32)
        for i, d in enumerate(prims):
            p[i] = [d["exponent"], *d["center"], d.get("coefficient", 1.0), float(bool(d.get("normalized", True)))]
            a[i] = d.get("angular", (0, 0, 0))
        out = np.zeros(2)
        f = np.ascontiguousarray(field, dtype=np.float64)
        self.check(self.lib.giaohip_primitive_eri(self.handle, _p(p), _p(a, C.c_int), _p(f), _p(out)))
        return complex(out[0], out[1])


class DeviceBasis:
    """A basis uploaded to one context (giaohip_basis)."""

    def __init__(self, ctx: Context, centers, L, nprim, ncontr, exps, coefs):
        self.ctx = ctx
        self._centers = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1)
        self._L = np.ascontiguousarray(L, dtype=np.int32)
        self._nprim = np.ascontiguousarray(nprim, dtype=np.int32)
        self._ncontr = np.ascontiguousarray(ncontr, dtype=np.int32)
        self._exps = np.ascontiguousarray(exps, dtype=np.float64)
        self._coefs = np.ascontiguousarray(coefs, dtype=np.float64)
        self.nshell = int(self._L.size)
        self.ao_counts = ((self._L + 1) * (self._L + 2) // 2 * self._ncontr).astype(np.int64)
        self.nao = int(self.ao_counts.sum())
        handle = _vp()
        ctx.check(
            ctx.lib.giaohip_basis_create(
                ctx.handle, C.c_int(self.nshell), _p(self._centers), _p(self._L, C.c_int), _p(self._nprim, C.c_int),
                _p(self._ncontr, C.c_int), _p(self._exps), _p(self._coefs), C.byref(handle),
            )
        )
        self.handle = handle

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.giaohip_basis_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def quartet_sizes(self, quartets) -> np.ndarray:
        q = np.asarray(quartets, dtype=np.int64).reshape(-1, 4)
        if q.size and (q.min() < 0 or q.max() >= self.nshell):
            raise IndexError("ERI shell-quartet index is out of range")
        return np.prod(self.ao_counts[q], axis=1) if len(q) else np.zeros(0, np.int64)

    def eri_batch(self, quartets, field, out=None, device_out: int | None = None):
        """Packed complex128 blocks of the quartets (host array), or into device_out."""
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.uint32).reshape(-1, 4))
        f = np.ascontiguousarray(field, dtype=np.float64)
        ctx = self.ctx
        if device_out is not None:  # the library validates the indices itself
            ctx.check(
                ctx.lib.giaohip_eri_batch(
                    ctx.handle, self.handle, _p(q, C.c_uint32), C.c_size_t(len(q)), _p(f), None, _vp(device_out), 1
                )
            )
            return None
        if out is None:
            out = np.empty(int(self.quartet_sizes(q).sum()), dtype=np.complex128)
        ctx.check(
            ctx.lib.giaohip_eri_batch(
                ctx.handle, self.handle, _p(q, C.c_uint32), C.c_size_t(len(q)), _p(f), None,
                out.ctypes.data_as(_vp), 0,
            )
        )
        return out

    def eri_derivative_batch(self, quartets, field, out=None, device_out: int | None = None):
        """Packed (4, 3, na, nb, nc, nd) centre-derivative blocks of the
        quartets (host complex128 array), or into device_out."""
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.uint32).reshape(-1, 4))
        f = np.ascontiguousarray(field, dtype=np.float64)
        ctx = self.ctx
        ok = len(q) and int(q.max()) < self.nshell  # else the library reports the index
        total = 12 * int(np.prod(self.ao_counts[q.astype(np.int64)], axis=1).sum()) if ok else 0
        if device_out is not None:
            ctx.check(ctx.lib.giaohip_eri_derivative_batch(
                ctx.handle, self.handle, _p(q, C.c_uint32), C.c_size_t(len(q)), _p(f), None, _vp(device_out),
                C.c_uint64(total), 1))
            return None
        if out is None:
            out = np.empty(total, dtype=np.complex128)
        ctx.check(ctx.lib.giaohip_eri_derivative_batch(
            ctx.handle, self.handle, _p(q, C.c_uint32), C.c_size_t(len(q)), _p(f), None, out.ctypes.data_as(_vp),
            C.c_uint64(out.size), 0))
        return out

    def eri_derivative_pair_bounds(self, field) -> np.ndarray:
        """EriDerivativeSchwarzBounds values, nshell x nshell."""
        out = np.zeros((self.nshell, self.nshell))
        f = np.ascontiguousarray(field, dtype=np.float64)
        self.ctx.check(self.ctx.lib.giaohip_eri_derivative_pair_bounds(self.ctx.handle, self.handle, _p(f), _p(out)))
        return out

    def schwarz_bounds(self, field) -> np.ndarray:
        """Shell-pair Schwarz factors (EriSchwarzBounds values), nshell x nshell."""
        out = np.empty((self.nshell, self.nshell))
        f = np.ascontiguousarray(field, dtype=np.float64)
        self.ctx.check(self.ctx.lib.giaohip_schwarz_bounds(self.ctx.handle, self.handle, _p(f), _p(out)))
        return out

    def screen(self, quartets, bounds, threshold: float) -> np.ndarray:
        """Quartets kept by Schwarz screening, in input order (indices validated)."""
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.uint32).reshape(-1, 4))
        kept = np.empty_like(q)
        nkept = C.c_size_t()
        bp = None if bounds is None else np.ascontiguousarray(bounds, dtype=np.float64)
        self.ctx.check(
            self.ctx.lib.giaohip_screen_quartets(
                self.ctx.handle, self.handle, _p(q, C.c_uint32), C.c_size_t(len(q)),
                None if bp is None else _p(bp), C.c_double(threshold), _p(kept, C.c_uint32), C.byref(nkept),
            )
        )
        return kept[: nkept.value]

    def eri_tensor(self, field):
        n = self.nao
        out = np.empty((n, n, n, n), dtype=np.complex128)
        f = np.ascontiguousarray(field, dtype=np.float64)
        self.ctx.check(self.ctx.lib.giaohip_eri_tensor(self.ctx.handle, self.handle, _p(f), out.ctypes.data_as(_vp), 0))
        return out

    def nuclear_attraction(self, charges, centers, field):
        """(nao, nao) complex128 nuclear-attraction matrix (compute_nuclear_attraction_matrix)."""
        n = self.nao
        out = np.empty((n, n), dtype=np.complex128)
        z = np.ascontiguousarray(charges, dtype=np.float64).reshape(-1)
        c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        if len(c) != len(z):
            raise ValueError("one center per nuclear charge")
        f = np.ascontiguousarray(field, dtype=np.float64)
        self.ctx.check(self.ctx.lib.giaohip_nuclear_attraction(self.ctx.handle, self.handle, len(z), _p(z), _p(c),
                                                               _p(f), out.ctypes.data_as(_vp)))
        return out

    def cholesky_diagonal(self, pairs, field):
        n = self.nao
        p = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint64).reshape(-1, 2))
        out = np.empty((n, n))
        f = np.ascontiguousarray(field, dtype=np.float64)
        self.ctx.check(
            self.ctx.lib.giaohip_cholesky_diagonal(
                self.ctx.handle, self.handle, _p(p, C.c_uint64), C.c_size_t(len(p)), _p(f), _p(out)
            )
        )
        return out

    def cholesky_columns(self, pivots, field, real: bool = False):
        n = self.nao
        p = np.ascontiguousarray(np.asarray(pivots, dtype=np.uint64).reshape(-1, 2))
        out = np.empty((len(p), n * n), dtype=np.float64 if real else np.complex128)
        f = np.ascontiguousarray(field, dtype=np.float64)
        self.ctx.check(
            self.ctx.lib.giaohip_cholesky_columns(
                self.ctx.handle, self.handle, _p(p, C.c_uint64), C.c_size_t(len(p)), _p(f),
                out.ctypes.data_as(_dp), C.c_int(int(real)),
            )
        )
        return out


class DeviceCholesky:
    """Factors of a device-resident Cholesky run (giaohip_cholesky)."""

    def __init__(self, basis: "DeviceBasis", handle):
        self.basis = basis
        self.ctx = basis.ctx
        self.handle = handle
        self._refresh()

    def _refresh(self):
        rank = C.c_uint64()
        res = C.c_double()
        tol = C.c_double()
        conv = C.c_int()
        col = C.c_double()
        upd = C.c_double()
        dia = C.c_double()
        self.ctx.check(
            self.ctx.lib.giaohip_cholesky_info(self.handle, C.byref(rank), C.byref(res), C.byref(tol), C.byref(conv),
                                               C.byref(col), C.byref(upd), C.byref(dia))
        )
        self.rank = int(rank.value)
        self.max_residual = float(res.value)
        self.tolerance = float(tol.value)
        self.converged = bool(conv.value)
        self.timings_ms = {"column": col.value, "update": upd.value, "diagonal": dia.value}
        blocks = C.c_uint64()
        rows = C.c_uint64()
        bms = C.c_double()
        proj = C.c_uint64()
        crows = C.c_uint64()
        self.ctx.check(self.ctx.lib.giaohip_cholesky_blocks(self.handle, C.byref(blocks), C.byref(rows), C.byref(bms),
                                                            C.byref(proj), C.byref(crows)))
        self.blocks = {"blocks": int(blocks.value), "factor_rows_read": int(rows.value), "block_ms": bms.value,
                       "projected": int(proj.value), "candidate_rows": int(crows.value)}
        gl, gi, gu = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.lib.giaohip_cholesky_graphs(self.handle, C.byref(gl), C.byref(gi), C.byref(gu)))
        self.graphs = {"launches": int(gl.value), "instantiated": int(gi.value), "updated": int(gu.value)}

    def remeasure(self) -> float:
        """Final residual max(D - sum |L|^2), all diagonal blocks measured again (cholesky.py:283-288)."""
        out = C.c_double()
        self.ctx.check(self.ctx.lib.giaohip_cholesky_remeasure(self.ctx.handle, self.handle, C.byref(out)))
        self._refresh()
        return float(out.value)

    def resume(self, stop_after: int = 0):
        """Continue the run until converged or the rank reaches stop_after (0: no limit)."""
        self.ctx.check(self.ctx.lib.giaohip_cholesky_resume(self.ctx.handle, self.handle, C.c_uint64(int(stop_after))))
        self._refresh()
        return self

    def pivots(self) -> np.ndarray:
        out = np.zeros((self.rank, 2), dtype=np.int64)
        self.ctx.check(self.ctx.lib.giaohip_cholesky_pivots(self.handle, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def vectors(self, first: int = 0, count: int | None = None) -> np.ndarray:
        n = self.basis.nao
        count = self.rank - first if count is None else count
        out = np.empty((count, n, n), dtype=np.complex128)
        self.ctx.check(
            self.ctx.lib.giaohip_cholesky_vectors(self.ctx.handle, self.handle, C.c_uint64(first), C.c_uint64(count),
                                                  out.ctypes.data_as(_dp))
        )
        return out

    def reconstruct(self, mu=(0, None), nu=(0, None), lam=(0, None), sig=(0, None)) -> np.ndarray:
        """(mu nu|lam sig) ~ sum_P L_P[mu,nu] conj(L_P[sig,lam]) on AO ranges (start, stop), on the device."""
        n = self.basis.nao
        r = []
        for a, b in (mu, nu, lam, sig):
            r += [int(a), n if b is None else int(b)]
        rr = np.ascontiguousarray(r, dtype=np.int64)
        shape = tuple(max(0, r[2 * k + 1] - r[2 * k]) for k in range(4))
        out = np.zeros(shape, dtype=np.complex128)
        self.ctx.check(self.ctx.lib.giaohip_cholesky_reconstruct(self.ctx.handle, self.handle,
                                                                 rr.ctypes.data_as(C.POINTER(C.c_int64)),
                                                                 out.ctypes.data_as(_dp)))
        return out

    def coulomb(self, density) -> np.ndarray:
        """J[mu,nu] = sum_{lam,sig} (mu nu|lam sig) D[lam,sig] from the device factors."""
        n = self.basis.nao
        d = np.ascontiguousarray(density, dtype=np.complex128).reshape(n, n)
        out = np.empty((n, n), dtype=np.complex128)
        self.ctx.check(self.ctx.lib.giaohip_cholesky_coulomb(self.ctx.handle, self.handle, d.ctypes.data_as(_dp),
                                                             out.ctypes.data_as(_dp)))
        return out

    def exchange(self, density) -> np.ndarray:
        """K[mu,nu] = sum_{lam,sig} (mu lam|sig nu) D[lam,sig] from the device factors."""
        n = self.basis.nao
        d = np.ascontiguousarray(density, dtype=np.complex128).reshape(n, n)
        out = np.empty((n, n), dtype=np.complex128)
        self.ctx.check(self.ctx.lib.giaohip_cholesky_exchange(self.ctx.handle, self.handle, d.ctypes.data_as(_dp),
                                                              out.ctypes.data_as(_dp)))
        return out

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.giaohip_cholesky_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cholesky_run(basis: "DeviceBasis", field, threshold: float, max_rank: int = 0, stop_after: int = 0):
    handle = _vp()
    f = np.ascontiguousarray(field, dtype=np.float64)
    basis.ctx.check(
        basis.ctx.lib.giaohip_cholesky_run(basis.ctx.handle, basis.handle, _p(f), C.c_double(float(threshold)),
                                           C.c_uint64(int(max_rank)), C.c_uint64(int(stop_after)), C.byref(handle))
    )
    return DeviceCholesky(basis, handle)


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    with _lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
        return _default_ctx
