"""Memory-bounded ERI interfaces of GIAOINT's Python package, over the HIP library.

Mirrors python/giao_integrals/eri.py:27-208: EriBatch, eri_batches, eri and
eri_schwarz_bounds, with the same arguments, results and error messages.
"""

from __future__ import annotations

from dataclasses import dataclass
from itertools import product
from math import prod

import numpy as np

from .core import (Basis, MagneticField, _eri_batch, _eri_derivative_batch, _eri_full, _EriDerivativeSchwarzBounds,
                   _EriSchwarzBounds, _threads)


@dataclass(frozen=True)
class EriBatch:
    """Packed shell-quartet blocks; values[offsets[i]:offsets[i+1]] reshaped to
    shapes[i] is the C-contiguous block of quartets[i]."""

    quartets: np.ndarray
    shapes: np.ndarray
    offsets: np.ndarray
    values: np.ndarray
    requested_count: int = 0
    screened_count: int = 0

    def block(self, index: int) -> np.ndarray:
        start, stop = int(self.offsets[index]), int(self.offsets[index + 1])
        return self.values[start:stop].reshape(tuple(self.shapes[index]))


def _quartets(shell_count: int, quartets):
    """Canonical representatives when quartets is None (eri.py:50-59), else
    the given quartets after the reference's checks (eri.py:61-79)."""
    if quartets is None:
        for q in product(range(shell_count), repeat=4):
            a, b, c, d = q
            if min(q, (c, d, a, b), (b, a, d, c), (d, c, b, a)) == q:
                yield q
        return
    for value in quartets:
        if len(value) != 4:
            raise ValueError("each shell quartet must contain four indices")
        if any(isinstance(i, (bool, np.bool_)) for i in value):
            raise ValueError("shell-quartet indices must be integers")
        try:
            q = tuple(int(i) for i in value)
        except (TypeError, ValueError, OverflowError) as error:
            raise ValueError("shell-quartet indices must be integers") from error
        if any(x != y for x, y in zip(q, value)):
            raise ValueError("shell-quartet indices must be integers")
        if any(i < 0 or i >= shell_count for i in q):
            raise IndexError("shell-quartet index is out of range")
        yield q


def _screening(value) -> float:
    v = float(value)
    if not np.isfinite(v) or v < 0.0:
        raise ValueError("screening_threshold must be finite and non-negative")
    return v


def eri_batches(basis: Basis, *, field: MagneticField | None = None, quartets=None, target_bytes: int = 64 << 20,
                screening_threshold: float = 0.0, threads: int | None = None):
    """Yield packed batches of about target_bytes (eri.py:82-145)."""
    if isinstance(target_bytes, (bool, np.bool_)) or not isinstance(target_bytes, (int, np.integer)) \
            or target_bytes <= 0:
        raise ValueError("target_bytes must be a positive integer")
    t = _screening(screening_threshold)
    _threads(threads)
    shells = basis.shells
    bounds = _EriSchwarzBounds(basis, field=field) if t > 0.0 else None
    pending: list = []
    size = 0

    def emit():
        q, shapes, offsets, values, requested, screened = _eri_batch(
            basis, pending, field=field, screening_threshold=t, bounds=bounds)
        return EriBatch(q, shapes, offsets, values, int(requested), int(screened))

    for q in _quartets(len(shells), quartets):
        nbytes = prod(shells[i].ao_count for i in q) * 16
        if pending and size + nbytes > target_bytes:
            yield emit()
            pending = []
            size = 0
        pending.append(q)
        size += nbytes
    if pending:
        yield emit()


def eri(basis: Basis, *, field: MagneticField | None = None, storage: str = "blocks", max_bytes: int | None = None,
        target_bytes: int = 64 << 20, screening_threshold: float = 0.0, threads: int | None = None):
    """Streamed blocks, or the full tensor under an explicit byte cap (eri.py:148-201)."""
    t = _screening(screening_threshold)
    _threads(threads)
    if storage == "blocks":
        if max_bytes is not None:
            raise ValueError("max_bytes is only valid with storage='full'")
        return eri_batches(basis, field=field, target_bytes=target_bytes, screening_threshold=t, threads=threads)
    if storage != "full":
        raise ValueError("storage must be 'blocks' or 'full'")
    if max_bytes is None:
        raise ValueError("storage='full' requires an explicit max_bytes limit")
    if isinstance(max_bytes, (bool, np.bool_)) or not isinstance(max_bytes, (int, np.integer)) or max_bytes < 0:
        raise ValueError("max_bytes must be a non-negative integer")
    required = basis.ao_count**4 * 16
    if required > max_bytes:
        raise MemoryError(f"full ERI tensor requires {required} bytes, exceeding max_bytes={max_bytes}")
    return _eri_full(basis, field=field, screening_threshold=t, threads=threads)


def eri_schwarz_bounds(basis: Basis, *, field: MagneticField | None = None) -> np.ndarray:
    """Shell-pair Schwarz factors (eri.py:204-208)."""
    return np.asarray(_EriSchwarzBounds(basis, field=field).values, dtype=np.float64)


@dataclass(frozen=True)
class EriDerivativeBatch(EriBatch):
    """Packed local centre derivatives, blocks (4, 3, na, nb, nc, nd)
    (eri.py:212-217): centre slots in quartet order, independent even for
    repeated shells; block() is a view of values; counts include screened
    requests."""


def eri_derivative_batches(basis: Basis, *, field: MagneticField | None = None, quartets=None,
                           target_bytes: int = 64 << 20, screening_threshold: float = 0.0, threads: int | None = None):
    """Stream ERI centre derivatives with the London phase response, at
    fixed B and gauge origin, without a full derivative AO tensor
    (eri.py:220-288). Canonical quartets when quartets is None, else the given
    ones in order; about target_bytes of packed values per batch (one
    oversized quartet alone). A positive screening_threshold drops quartets
    whose derivative-density Schwarz bound is below it."""
    if isinstance(target_bytes, (bool, np.bool_)) or not isinstance(target_bytes, (int, np.integer)) \
            or target_bytes <= 0:
        raise ValueError("target_bytes must be a positive integer")
    threshold = float(screening_threshold)
    if not np.isfinite(threshold) or threshold < 0:
        raise ValueError("screening_threshold must be finite and non-negative")
    _threads(threads)
    bounds = _EriDerivativeSchwarzBounds(basis, field=field, threads=threads) if threshold > 0 else None
    shells = basis.shells
    pending: list = []
    size = 0

    def emit():
        q, shapes, offsets, values, requested, screened = _eri_derivative_batch(
            basis, pending, field=field, screening_threshold=threshold, threads=threads, bounds=bounds)
        return EriDerivativeBatch(q, shapes, offsets, values, requested, screened)

    for q in _quartets(len(shells), quartets):
        block = 12 * prod(shells[i].ao_count for i in q) * 16
        if pending and size + block > target_bytes:
            yield emit()
            pending.clear()
            size = 0
        pending.append(q)
        size += block
    if pending:
        yield emit()
