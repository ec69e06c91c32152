"""Pivoted Cholesky of the ordered-pair Coulomb matrix, over the HIP library.

Mirrors python/giao_integrals/cholesky.py:16-297 (CholeskyERI, cholesky_eri).
Every run stays on the device end to end (giaohip_cholesky_run). Zero-field
and one-centre runs use real arithmetic with symmetrized columns; they, and
complex runs on bases with one contraction per shell, pivot on the global
maximum (cholesky.py:175-247). Complex runs on bases with general
contractions follow the reference's candidate-block policy (contraction
variants of the pivot's component pair, the pivot taken from the latest
block when its largest residual is at least 1% of the maximum, and the
raw-column cache whose membership decides when a new block forms,
cholesky.py:193-262); the library keeps that bookkeeping on the host and
computes every column when its pivot is taken.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native
from .core import Basis, MagneticField, _EriSchwarzBounds, _field6, _threads


@dataclass(frozen=True)
class CholeskyERI:
    """Coulomb factors: (mu nu|lam sig) ~ sum_P L[P,mu,nu] conj(L[P,sig,lam]).

    vectors complex128 (rank, nao, nao); pivots int64 (rank, 2) AO pairs;
    threshold as requested; max_residual_diagonal the measured largest
    remaining diagonal of W[(mu,nu),(lam,sig)] = (nu mu|lam sig).
    """

    vectors: np.ndarray
    pivots: np.ndarray
    threshold: float
    max_residual_diagonal: float

    def reconstruct(self, slices=None) -> np.ndarray:
        """Four AO slices (mu, nu, lam, sig) of the reconstructed ERIs (cholesky.py:39-57)."""
        if slices is None:
            slices = (slice(None),) * 4
        if len(slices) != 4 or not all(isinstance(s, slice) for s in slices):
            raise ValueError("slices must contain four slices")
        a, b, c, d = slices
        left = self.vectors[:, a, b]
        right = self.vectors[:, d, c].conj()
        m, n = left.shape[1:], right.shape[1:]
        out = left.reshape(len(left), -1).T @ right.reshape(len(right), -1)
        # right[P, s, l] -> (mu, nu, l, s) order
        return out.reshape(*m, *n).transpose(0, 1, 3, 2)


def _nonnegative(name: str, value) -> float:
    v = float(value)
    if not np.isfinite(v) or v < 0:
        raise ValueError(f"{name} must be finite and non-negative")
    return v


def cholesky_eri(basis: Basis, *, field: MagneticField | None = None, threshold: float = 1e-8,
                 screening_threshold: float = 0.0, threads: int | None = None, max_bytes: int | None = None,
                 pivot_batch_size: int = 32) -> CholeskyERI:
    """Factor the ERIs without forming the AO tensor (cholesky.py:60-297).

    Stops when the largest residual diagonal is <= max(threshold,
    64 eps max(initial diagonal)). With screening_threshold > 0 the
    reference starts a shell pair whose squared Schwarz bound is at most
    min(threshold, screening_threshold) from that bound, measures it after
    the loop and reports max(D - sum |L|^2) as the final residual
    (cholesky.py:138-154,283-288). Such entries are never pivots or
    candidates (their residual is below the tolerance), so the device run,
    which measures every diagonal block up front, takes the same pivots; when
    any pair would have been skipped, the final residual is re-measured the
    same way on the device. max_bytes caps the returned vectors and pivot
    indices (MemoryError beyond it).
    """
    threshold = _nonnegative("threshold", threshold)
    screening_threshold = _nonnegative("screening_threshold", screening_threshold)
    _threads(threads)
    if isinstance(pivot_batch_size, (bool, np.bool_)) or not isinstance(pivot_batch_size, (int, np.integer)) \
            or pivot_batch_size < 1:
        raise ValueError("pivot_batch_size must be a positive integer")
    if max_bytes is not None and (isinstance(max_bytes, (bool, np.bool_))
                                  or not isinstance(max_bytes, (int, np.integer)) or max_bytes < 0):
        raise ValueError("max_bytes must be a non-negative integer or None")
    n = basis.ao_count
    row_bytes = n * n * 16 + 16
    limit = n * n if max_bytes is None else min(n * n, int(max_bytes) // row_bytes)
    dev = basis._dev()
    dev.ctx.set_cholesky_pivot_batch(int(pivot_batch_size))
    f = _field6(field)
    screened = False
    if screening_threshold and len(basis.shells):
        # cholesky.py:128-154: bounds at the integral field (zero on one centre)
        centers = [s.center for s in basis.shells]
        integral_field = None if all(c == centers[0] for c in centers) else field
        bounds = _EriSchwarzBounds(basis, field=integral_field).values ** 2
        upper = np.triu(bounds <= min(threshold, screening_threshold))
        screened = bool(np.any(upper))

    def over(rank):
        return MemoryError(f"Cholesky ERI output requires at least {(rank + 1) * row_bytes} bytes, "
                           f"exceeding max_bytes={max_bytes}")

    if limit == 0:
        probe = _native.cholesky_run(dev, f, threshold, 0, 1)
        if probe.rank:
            raise over(0)
        run = probe
    else:
        try:
            run = _native.cholesky_run(dev, f, threshold, limit if max_bytes is not None else 0)
        except MemoryError:
            raise over(limit) from None
    if screened:
        run.remeasure()
    vectors = run.vectors() if run.rank else np.zeros((0, n, n), dtype=np.complex128)
    return CholeskyERI(vectors, run.pivots().astype(np.int64).reshape(-1, 2), threshold, float(run.max_residual))
