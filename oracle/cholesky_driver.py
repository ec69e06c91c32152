"""Checker restatement of the reference's Cholesky driver (TEST INFRASTRUCTURE).

Restates python/giao_integrals/cholesky.py:60-297 in numpy with the
selected-integral calls injected: ``diagonal(pairs) -> (n, n)`` and
``columns(pivots) -> (len(pivots), n*n)`` (complex unless ``real``). With
the C oracle's (or the reference-built library's) diagonal and columns this
reproduces the reference's pivot policy: strict global argmax (first maximum
in flat mu*n+nu order), the local-block override for complex generally
contracted factors, the candidate sets and the bounded raw-column cache.
Schwarz-screened diagonals (screening_threshold > 0) follow
cholesky.py:138-154,283-288: a shell pair whose squared Schwarz bound is at
most min(threshold, screening_threshold) starts from that bound instead of
its measured diagonal and is measured once the loop ends, when the final
residual becomes D - sum |L|^2 (``bounds`` gives the squared bounds).
The general-contraction policy is pinned by the reference's recorded runs on
water cc-pVTZ (ranks 656 / 668 / 687 and their residuals,
tests/test_oracle_ccpvtz.py); its factorization also by the dense
factorization in the same pivot order.
"""

from __future__ import annotations

from collections import OrderedDict

import numpy as np


def cholesky_eri(ao_offsets, shells, diagonal, columns, *, real: bool, single_center: bool,
                 threshold: float = 1e-8, max_bytes=None, pivot_batch_size: int = 32, max_pivots=None,
                 history=None, follow=None, tie_rtol: float = 1e-12, stats=None, bounds=None,
                 screening_threshold: float = 0.0):
    """shells: list of (cartesian_count, contraction_count, primitive_count).

    Returns (vectors (rank, n, n) complex128, pivots (rank, 2), max residual,
    tolerance). ``max_pivots`` truncates the loop (prefix timing only);
    ``history`` (a list) receives a copy of the residual before every step.

    ``follow`` replays another implementation's pivot sequence: at every
    step the followed pivot must be the policy's choice up to rounding, i.e.
    its residual within ``tie_rtol * max(initial diagonal)`` of the largest
    residual of the pool the policy takes the pivot from (everything, or the
    latest candidate block when the local override applies; residuals
    D - sum |L|^2 carry absolute rounding of order eps * max D per update);
    otherwise AssertionError. ``stats`` (a dict) receives the number
    of steps where the followed pivot differs from this run's own argmax.
    """
    offsets = np.asarray(ao_offsets, dtype=np.int64)
    ncart = np.array([s[0] for s in shells])
    ncontr = np.array([s[1] for s in shells])
    nprim = np.array([s[2] for s in shells])
    ao = ncart * ncontr
    n = int(ao.sum())
    ns = len(shells)
    spans = [slice(int(o), int(o + a)) for o, a in zip(offsets, ao)]
    owners = np.repeat(np.arange(ns), ao)
    diag = np.zeros((n, n))
    skipped, measured_pairs = [], []
    for a in range(ns):
        for b in range(a, ns):
            # cholesky.py:138-154: squared Schwarz bounds stand in for pairs
            # already below min(threshold, screening_threshold)
            if screening_threshold and bounds is not None and bounds[a, b] <= min(threshold, screening_threshold):
                diag[spans[a], spans[b]] = bounds[a, b]
                diag[spans[b], spans[a]] = bounds[a, b]
                skipped.append((a, b))
            else:
                measured_pairs.append((a, b))

    def read_diagonal(pairs):
        if not pairs:
            return
        measured = diagonal(pairs)
        for a, b in pairs:
            diag[spans[a], spans[b]] = measured[spans[a], spans[b]]
            diag[spans[b], spans[a]] = measured[spans[b], spans[a]]

    read_diagonal(measured_pairs)
    residual = diag.ravel().copy()
    scale = float(np.max(residual, initial=0.0))
    floor = 64 * np.finfo(float).eps * scale
    tolerance = max(threshold, floor)
    dtype = np.float64 if real else np.complex128
    row_bytes = n * n * 16 + 16
    limit = n * n if max_bytes is None else min(n * n, max_bytes // row_bytes)
    factors = np.empty((0, n * n), dtype=dtype)
    pivots = []
    reuse = not single_center or bool(np.any(ncontr > 1) or np.any(nprim > 1))
    candidate_count = 32 if reuse else 8
    raw_cache = OrderedDict()
    block = np.empty(0, dtype=int)
    candidates = np.empty(0, dtype=int)
    cols = np.empty((0, n * n), dtype=dtype)
    while float(np.max(residual, initial=0.0)) > tolerance:
        if max_pivots is not None and len(pivots) >= max_pivots:
            break
        if history is not None:
            history.append(residual.copy())
        pivot = int(np.argmax(residual))
        overridden = False
        if reuse and not real and pivot_batch_size > 1 and len(block):
            local = int(block[np.argmax(residual[block])])
            if residual[local] > tolerance and residual[local] >= 0.01 * residual[pivot]:
                pivot = local
                overridden = True
                if stats is not None:
                    stats["overrides"] = stats.get("overrides", 0) + 1
        if follow is not None:
            # the followed pivot must be this policy's choice up to rounding:
            # a maximum of the pool the policy chose from (the block when it
            # overrode the global maximum, else everything) within tie_rtol
            k = len(pivots)
            if k >= len(follow):
                raise AssertionError(f"followed run stopped after {k} pivots; residual {residual[pivot]:.3e} > {tolerance:.3e}")
            fp = int(follow[k][0]) * n + int(follow[k][1])
            top = residual[pivot]
            in_pool = (not overridden) or fp in set(int(x) for x in block)
            if not (in_pool and residual[fp] >= top - tie_rtol * scale):
                raise AssertionError(
                    f"step {k}: followed pivot {divmod(fp, n)} residual {residual[fp]!r} is not the policy's "
                    f"choice ({divmod(pivot, n)} has {top!r})"
                )
            if fp != pivot and stats is not None:
                stats["tie_flips"] = stats.get("tie_flips", 0) + 1
            pivot = fp
        rank = len(pivots)
        if rank >= limit:
            raise MemoryError(f"Cholesky ERI output requires at least {(rank + 1) * row_bytes} bytes")
        if rank == len(factors):
            cap = min(limit, max(8, 2 * len(factors)))
            grown = np.empty((cap, n * n), dtype=dtype)
            grown[:rank] = factors[:rank]
            factors = grown
        if reuse and pivot in raw_cache:
            candidates = np.array([pivot])
            cols = raw_cache.pop(pivot)[None, :]
        match = np.flatnonzero(candidates == pivot)
        if not len(match):
            mu, nu = divmod(pivot, n)
            a, b = int(owners[mu]), int(owners[nu])
            if reuse:
                left = offsets[a] + (mu - offsets[a]) % ncart[a] + ncart[a] * np.arange(ncontr[a])
                right = offsets[b] + (nu - offsets[b]) % ncart[b] + ncart[b] * np.arange(ncontr[b])
                idx = (left[:, None] * n + right).ravel()
            else:
                idx = np.arange(n * n).reshape(n, n)[spans[a], spans[b]].ravel()
            idx = idx[residual[idx] > tolerance]
            candidates = idx[np.argsort(-residual[idx], kind="stable")[:candidate_count]]
            if pivot not in candidates:
                candidates[-1] = pivot
            block = candidates[: min(32, pivot_batch_size)]
            cmu, cnu = np.divmod(candidates, n)
            cols = columns(list(zip(cmu.tolist(), cnu.tolist())))
            if reuse:
                for i, c in enumerate(candidates):
                    if c != pivot:
                        raw_cache[int(c)] = cols[i].copy()
                cap = min(1024, max(8, rank + 8))
                while len(raw_cache) > cap:
                    raw_cache.popitem(last=False)
            match = np.flatnonzero(candidates == pivot)
        if reuse:
            raw_cache.pop(pivot, None)
        column = cols[int(match[0])].copy()
        if rank:
            column -= factors[:rank, pivot].conj() @ factors[:rank]
        if real:
            sq = column.reshape(n, n)
            column = ((sq + sq.T) * 0.5).ravel()
        column /= np.sqrt(residual[pivot])
        factors[rank] = column
        pivots.append(divmod(pivot, n))
        residual -= np.abs(column) ** 2
        if float(np.min(residual, initial=0.0)) < -max(floor * 16, 1e-12):
            raise RuntimeError("Coulomb Cholesky residual lost positive semidefiniteness")
        np.maximum(residual, 0, out=residual)
        residual[pivot] = 0
    if follow is not None and len(follow) != len(pivots) and max_pivots is None:
        raise AssertionError(f"followed run has {len(follow)} pivots, this run stopped after {len(pivots)}")
    if skipped:
        # cholesky.py:283-288: measure the skipped pairs, then the final
        # residual is D - sum |L|^2 (not clamped, pivots not zeroed)
        read_diagonal(skipped)
        residual = diag.ravel() - np.sum(np.abs(factors[: len(pivots)]) ** 2, axis=0)
        if stats is not None:
            stats["skipped_pairs"] = len(skipped)
    vectors = np.array(factors[: len(pivots)], dtype=np.complex128).reshape(len(pivots), n, n)
    return vectors, np.asarray(pivots, dtype=np.int64).reshape(-1, 2), float(np.max(residual, initial=0.0)), tolerance
