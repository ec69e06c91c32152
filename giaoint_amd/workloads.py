"""Deterministic synthetic inputs of the five measured workloads.

Every generator uses numpy's PCG64 with a fixed seed and returns raw
(un-normalized) contraction coefficients; normalization happens where a basis
is built, exactly as ``giao.Shell(..., normalize=True)`` does.

* ``water_ccpvdz``   water (O 0,0,0.1173 A; H 0,+-0.7572,-0.4692 A) in
                     cc-pVDZ, 7 shells / 25 Cartesian AOs, B=(0,0,0.1).
* ``dddd_batch``     200k (dd|dd) quartets, 4 private shells each, centres
                     U([0,6]^3) bohr, exponents 10^U(-1,4), 3 primitives,
                     coefficients N(0,1), one field drawn uniformly from
                     |B|<=0.5 (seed 2).
* ``mixed_batch``    2M quartets over classes l=0..3 with P ~ (1+sum l)^-2,
                     contraction degree U{1..8} per shell (seed 3).
* ``precision_batches`` 50k (gg|gg) and 200k (ff|dd), exponents
                     10^U(log10 .05, log10 5e3), degree 2 (seed 4).
* ``cholesky_atoms`` 30 atoms in a 14-bohr box (min spacing 1.8), each with
                     uncontracted even-tempered 4s3p2d1f (alpha_i=0.05*3^i),
                     B=(0.05,0,0.1) (seed 5).

Quartets of the random batches are redrawn while any of their primitive
quartets has a Boys argument outside the reference's verified domain
(|T|>160 outside the positive asymptotic sector), where the reference
raises BoysNumericalError (boys.cpp:262-270): such inputs are not evaluable
by the path being replaced.
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

BOHR_IN_ANGSTROM = 0.529177210544  # molecule.py:26
DATA = Path(__file__).resolve().parent / "data"


@dataclass
class ShellArrays:
    """Flat shell description (C-ABI layout), raw coefficients."""

    centers: np.ndarray  # (ns, 3)
    L: np.ndarray  # (ns,)
    nprim: np.ndarray
    ncontr: np.ndarray
    exps: np.ndarray
    coefs: np.ndarray  # raw, per shell [contraction][primitive]
    labels: list = field(default_factory=list)

    @property
    def nshell(self) -> int:
        return int(self.L.size)

    def shell_list(self):
        """[(center, L, exps, coefs(ncontr, nprim))] for the checkers."""
        out = []
        po = co = 0
        for s in range(self.nshell):
            np_, nc = int(self.nprim[s]), int(self.ncontr[s])
            out.append(
                (self.centers[s], int(self.L[s]), self.exps[po : po + np_], self.coefs[co : co + np_ * nc].reshape(nc, np_))
            )
            po += np_
            co += np_ * nc
        return out

    def subset(self, shell_ids) -> "ShellArrays":
        ids = np.asarray(shell_ids, dtype=np.int64)
        po = np.concatenate([[0], np.cumsum(self.nprim)])
        co = np.concatenate([[0], np.cumsum(self.nprim * self.ncontr)])
        exps = np.concatenate([self.exps[po[i] : po[i + 1]] for i in ids]) if len(ids) else np.zeros(0)
        coefs = np.concatenate([self.coefs[co[i] : co[i + 1]] for i in ids]) if len(ids) else np.zeros(0)
        return ShellArrays(self.centers[ids], self.L[ids], self.nprim[ids], self.ncontr[ids], exps, coefs)


def _from_shells(shells) -> ShellArrays:
    centers = np.array([s[0] for s in shells], dtype=np.float64).reshape(-1, 3)
    L = np.array([s[1] for s in shells], dtype=np.int32)
    nprim = np.array([len(s[2]) for s in shells], dtype=np.int32)
    ncontr = np.array([np.atleast_2d(s[3]).shape[0] for s in shells], dtype=np.int32)
    exps = np.concatenate([np.asarray(s[2], float) for s in shells])
    coefs = np.concatenate([np.atleast_2d(np.asarray(s[3], float)).reshape(-1) for s in shells])
    return ShellArrays(centers, L, nprim, ncontr, exps, coefs)


# ---------------------------------------------------------------------------
# Water cc-pVDZ
# ---------------------------------------------------------------------------
WATER_ANGSTROM = [("O", (0.0, 0.0, 0.1173)), ("H", (0.0, 0.7572, -0.4692)), ("H", (0.0, -0.7572, -0.4692))]


def bse_shells(atoms_bohr, data) -> list:
    """Shells from BSE JSON data in the Molecule front end's order (atom
    order, then stable increasing L; a general contraction stays one shell),
    as (centre, L, exponents, raw coefficient rows) tuples."""
    from .molecule import _element_shells

    shells = []
    for sym, pos in atoms_bohr:
        for l, exps, rows in _element_shells(data["elements"], sym):
            shells.append((np.array(pos, float), int(l), np.array(exps), np.array(rows)))
    return shells


def water_ccpvdz():
    data = json.loads((DATA / "cc-pvdz-h-o.json").read_text())
    atoms = [(s, tuple(x / BOHR_IN_ANGSTROM for x in p)) for s, p in WATER_ANGSTROM]
    shells = bse_shells(atoms, data)
    return _from_shells(shells), (0.0, 0.0, 0.1), (0.0, 0.0, 0.0)


# ---------------------------------------------------------------------------
# Random shell-quartet batches
# ---------------------------------------------------------------------------
def _boys_domain_ok(ea, A, ka, eb, B, kb, ec, Cc, kc, ed, D, kd, L):
    """True where every primitive quartet of a quartet has a Boys argument in
    the verified domain. Arrays are (n, K) exponents and (n, 3) centres."""
    def pair(e1, c1, k1, e2, c2, k2):
        p = e1[:, :, None] + e2[:, None, :]  # (n, K1, K2)
        P = (e1[:, :, None, None] * c1[:, None, None, :] + e2[:, None, :, None] * c2[:, None, None, :]) / p[..., None]
        q = (k2 - k1)[:, None, None, :]
        Pc = P - 0.5j * q / p[..., None]
        n = p.shape[0]
        return p.reshape(n, -1), Pc.reshape(n, -1, 3)

    p, Pc = pair(ea, A, ka, eb, B, kb)
    q, Qc = pair(ec, Cc, kc, ed, D, kd)
    rho = p[:, :, None] * q[:, None, :] / (p[:, :, None] + q[:, None, :])
    X = Pc[:, :, None, :] - Qc[:, None, :, :]
    T = rho * np.sum(X * X, axis=-1)
    mag = np.abs(T)
    asym = (T.real >= 2.0 * L[:, None, None] + 60.0) & (np.abs(T.imag) <= 0.5 * T.real)
    ok = (mag <= 160.0) | asym
    return ok.reshape(ok.shape[0], -1).all(axis=1)


def _wave(B, c):
    Bx, By, Bz = B
    return 0.5 * np.stack([By * c[:, 2] - Bz * c[:, 1], Bz * c[:, 0] - Bx * c[:, 2], Bx * c[:, 1] - By * c[:, 0]], axis=1)


def _ball(rng, radius):
    while True:
        v = rng.uniform(-radius, radius, 3)
        if v @ v <= radius * radius:
            return tuple(float(x) for x in v)


@dataclass
class QuartetBatch:
    shells: ShellArrays
    quartets: np.ndarray  # (n, 4) uint32
    B: tuple
    origin: tuple
    seed: int
    description: str


def random_quartets(classes, degrees, *, seed, box=6.0, log_exp=(-1.0, 4.0), B=None, name="") -> QuartetBatch:
    """Quartets with four private shells each. classes: (n,4) ints; degrees: (n,4)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    classes = np.asarray(classes, dtype=np.int32)
    degrees = np.asarray(degrees, dtype=np.int32)
    n = len(classes)
    if B is None:
        B = _ball(rng, 0.5)
    kmax = int(degrees.max()) if n else 1
    centers = np.empty((n, 4, 3))
    exps = np.empty((n, 4, kmax))
    todo = np.arange(n)
    while len(todo):
        m = len(todo)
        centers[todo] = rng.uniform(0.0, box, (m, 4, 3))
        exps[todo] = 10.0 ** rng.uniform(log_exp[0], log_exp[1], (m, 4, kmax))
        # Boys-domain check per contraction-degree group (the check is per
        # quartet; grouping only lets it run vectorized)
        bad = []
        uniq, inv = np.unique(degrees[todo], axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        order = np.argsort(inv, kind="stable")
        edges = np.searchsorted(inv[order], np.arange(len(uniq) + 1))
        for g, deg in enumerate(uniq):
            sel = todo[order[edges[g]:edges[g + 1]]]
            L = classes[sel].sum(axis=1)
            c = centers[sel]
            k = [_wave(B, c[:, i]) for i in range(4)]
            step = max(1, 4096 // int(np.prod(deg)) * 16)
            for start in range(0, len(sel), step):
                s = slice(start, start + step)
                ok = _boys_domain_ok(
                    exps[sel[s], 0, : deg[0]], c[s, 0], k[0][s], exps[sel[s], 1, : deg[1]], c[s, 1], k[1][s],
                    exps[sel[s], 2, : deg[2]], c[s, 2], k[2][s], exps[sel[s], 3, : deg[3]], c[s, 3], k[3][s], L[s],
                )
                bad.extend(sel[s][~ok].tolist())
        todo = np.array(sorted(bad), dtype=np.int64)
    # the first degrees[i, j] exponents of every shell, shell by shell
    used = np.arange(kmax)[None, None, :] < degrees[:, :, None]
    coef_rng = np.random.Generator(np.random.PCG64(seed + 1000))
    flat_deg = degrees.reshape(-1)
    allc = coef_rng.normal(size=int(flat_deg.sum()))
    shells = ShellArrays(
        centers.reshape(-1, 3).copy(),
        classes.reshape(-1).copy(),
        flat_deg.copy(),
        np.ones(4 * n, dtype=np.int32),
        exps[used] if n else np.zeros(0),
        allc,
    )
    # guard against all-zero contractions (N(0,1) draws are never exactly 0)
    quartets = np.arange(4 * n, dtype=np.uint32).reshape(n, 4)
    return QuartetBatch(shells, quartets, tuple(B), (0.0, 0.0, 0.0), seed, name)


def dddd_batch(n: int = 200_000, seed: int = 2) -> QuartetBatch:
    return random_quartets(np.full((n, 4), 2), np.full((n, 4), 3), seed=seed, name=f"{n} (dd|dd) K=3^4")


def mixed_batch(n: int = 2_000_000, seed: int = 3) -> QuartetBatch:
    rng = np.random.Generator(np.random.PCG64(seed + 2000))
    cls = np.array(np.meshgrid(*[np.arange(4)] * 4, indexing="ij")).reshape(4, -1).T
    w = (1.0 + cls.sum(axis=1)) ** -2.0
    pick = rng.choice(len(cls), size=n, p=w / w.sum())
    deg = rng.integers(1, 9, size=(n, 4))
    return random_quartets(cls[pick], deg, seed=seed, name=f"{n} mixed l<=3, degree U{{1..8}}")


def precision_batches(seed: int = 4, n_gggg: int = 50_000, n_ffdd: int = 200_000):
    lo, hi = np.log10(0.05), np.log10(5e3)
    g = random_quartets(np.full((n_gggg, 4), 4), np.full((n_gggg, 4), 2), seed=seed, log_exp=(lo, hi),
                        name=f"{n_gggg} (gg|gg) K=2^4")
    f = random_quartets(np.tile([3, 3, 2, 2], (n_ffdd, 1)), np.full((n_ffdd, 4), 2), seed=seed + 10,
                        log_exp=(lo, hi), name=f"{n_ffdd} (ff|dd) K=2^4")
    return g, f


def cholesky_atoms(natoms: int = 30, seed: int = 5, box: float = 14.0, min_sep: float = 1.8):
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = []
    while len(pos) < natoms:
        c = rng.uniform(0.0, box, 3)
        if all(np.linalg.norm(c - p) >= min_sep for p in pos):
            pos.append(c)
    shells = []
    for c in pos:
        for l, count in ((0, 4), (1, 3), (2, 2), (3, 1)):
            for i in range(count):
                shells.append((c, l, np.array([0.05 * 3.0**i]), np.array([[1.0]])))
    return _from_shells(shells), (0.05, 0.0, 0.1), (0.0, 0.0, 0.0)


# ---------------------------------------------------------------------------
# Algorithmic FLOP convention (collapsed complex-E MD formulation)
# ---------------------------------------------------------------------------
def _nc(l):
    return (l + 1) * (l + 2) // 2


def _nh(L):
    return (L + 1) * (L + 2) * (L + 3) // 6


def _r0(L):
    return sum((d + 1) * (d + 2) // 2 * (L - d + 1) for d in range(L + 1))


def quartet_flop_parts(la, lb, lc, ld, ka, kb, kc, kd) -> tuple:
    """(R0 part, contraction part) of the FLOP convention for one shell quartet.

    R0 part: 12 per R0 entry per primitive quartet. Contraction part: 8 per
    complex MAC of the cheaper of the bra-deferred and ket-deferred dense
    Hermite-to-Cartesian contractions.
    """
    K = ka * kb * kc * kd
    L = la + lb + lc + ld
    nab, ncd = _nc(la) * _nc(lb), _nc(lc) * _nc(ld)
    hab, hcd = _nh(la + lb), _nh(lc + ld)
    a = K * 8 * hab * hcd * ncd + ka * kb * 8 * nab * hab * ncd
    b = K * 8 * hab * hcd * nab + kc * kd * 8 * nab * hcd * ncd
    return float(K * 12 * _r0(L)), float(min(a, b))


def quartet_flops(la, lb, lc, ld, ka, kb, kc, kd) -> float:
    """F = 12 K R0(L) + min(bra-deferred, ket-deferred) complex-MAC cost."""
    return sum(quartet_flop_parts(la, lb, lc, ld, ka, kb, kc, kd))


def _cart(l):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def plan_entries(l1: int, l2: int) -> int:
    """Nonzero complex Hermite products of a pair: sum over components of the box size."""
    n = 0
    for a in _cart(l1):
        for b in _cart(l2):
            n += (a[0] + b[0] + 1) * (a[1] + b[1] + 1) * (a[2] + b[2] + 1)
    return n


def quartet_executed_flops(la, lb, lc, ld, ka, kb, kc, kd) -> float:
    """FLOPs the batched transform executes (8 per complex MAC, structural zeros skipped).

    Ket step: every (bra pair, bra Hermite index) row meets every nonzero ket
    product of every ket pair; bra step: every nonzero bra product meets every
    ket component, once per bra pair. The pair with fewer components is the bra
    side. The R0 tables are counted as in the convention (12 per entry).
    """
    if _nc(la) * _nc(lb) > _nc(lc) * _nc(ld):
        la, lb, lc, ld, ka, kb, kc, kd = lc, ld, la, lb, kc, kd, ka, kb
    K = ka * kb * kc * kd
    ket = K * _nh(la + lb) * plan_entries(lc, ld)
    bra = ka * kb * plan_entries(la, lb) * _nc(lc) * _nc(ld)
    return float(K * 12 * _r0(la + lb + lc + ld) + 8 * (ket + bra))


def batch_executed_flops(shells: ShellArrays, quartets: np.ndarray) -> float:
    q = np.asarray(quartets, dtype=np.int64)
    l = shells.L[q]
    k = shells.nprim[q]
    keys, inv = np.unique(np.concatenate([l, k], axis=1), axis=0, return_inverse=True)
    counts = np.bincount(inv.reshape(-1), minlength=len(keys))
    return float(sum(c * quartet_executed_flops(*map(int, key)) for key, c in zip(keys, counts)))


def batch_flops(shells: ShellArrays, quartets: np.ndarray, part: str = "all") -> float:
    """Convention FLOPs of a batch; part "all", "r0" or "contract"."""
    q = np.asarray(quartets, dtype=np.int64)
    l = shells.L[q]
    k = shells.nprim[q]
    keys, inv = np.unique(np.concatenate([l, k], axis=1), axis=0, return_inverse=True)
    counts = np.bincount(inv.reshape(-1), minlength=len(keys))
    pick = {"all": lambda r, c: r + c, "r0": lambda r, c: r, "contract": lambda r, c: c}[part]
    return float(sum(n * pick(*quartet_flop_parts(*map(int, key))) for key, n in zip(keys, counts)))
