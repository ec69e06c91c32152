"""Autotuner for the launch parameters of the ERI and Cholesky kernels.

GIAOINT has no counterpart: its OpenMP path has no tile or launch parameters
(`src/eri.cpp:730-756` schedules shell quartets `dynamic,1`). On the GPU four
parameters are chosen at run time, by timing the caller's own workload on the
device in use:

* ``fast_waves`` -- workgroup size of the class-specialized transform
  (4..8 waves). A workgroup takes as many bra primitive pairs per pass as its
  rows hold, so this is the bra-pair tile of the (dd|dd), (dd|ff), (pp|pp),
  (ps|ps) and (ss|ss) kernels.
* ``eri_queue`` -- launch form of the generic transform: one workgroup per
  item and LDS-size bucket, a persistent device work queue per bucket, or one
  persistent launch over every class.
* ``chunk_mib`` -- work memory of one chunk of the batched path (chunks
  alternate between the producer and the contraction stream, so the size
  trades pipeline fill and drain against launch count).
* ``cholesky_block`` -- candidates per pivot block of the streamed Cholesky
  (projection tile width against passes over the factors).
* ``wave_max_l`` -- largest total angular momentum whose generic-transform
  items run on one-wavefront workgroups (four times the items in flight per
  CU, against fewer threads per item).
* ``mfma_min_l`` -- smallest total angular momentum whose classes (without a
  class-specialized kernel) run their contraction on the FP64 matrix cores
  (dense products) instead of the generic FP64 VALU kernel.
* ``small_max_l`` -- largest total angular momentum of the classes on the
  lane-per-primitive-quartet kernel (-1 turns it off).
* ``mfma_variant`` -- a compile-time choice of the matrix-core kernels: every
  (LAB, LCD) instance is built in four variants (GEMM step unroll 1 or 2,
  LDS row stride of its W and ket-product tiles padded by 16 bytes or not),
  and the tuner picks the variant that runs the caller's workload fastest.

Every candidate is checked before it can win: ERI candidates must reproduce
the default configuration's integrals on a sample of the workload (bitwise
for the launch forms, the chunk size and the one-wavefront workgroups, which
do not change any sum; within ``tolerance`` for the class-specialized
workgroup size, which changes the order of the bra-pair sums, and for the
matrix-core threshold, which changes the kernel), and
Cholesky candidates must give the same pivots and factors bitwise (the block
size is designed not to change them). Choices can be kept in a JSON cache
keyed by device and workload (``save`` / ``load``).
"""

from __future__ import annotations

import json
import os
import time
from dataclasses import asdict, dataclass, field
from pathlib import Path

import numpy as np

from . import _native

FAST_WAVES = (4, 5, 6, 7, 0)
ERI_QUEUE = ("none", "buckets", "single")
CHUNK_MIB = (1024, 2048, 4096, 8192, 16384)
CHOLESKY_BLOCK = (8, 16, 24, 32)
WAVE_MAX_L = (-1, 5, 6, 7, 8)
MFMA_MIN_L = (7, 9, 11, 14)
MFMA_VARIANT = (0, 1, 2, 3)
SMALL_MAX_L = (-1, 2)
DEFAULTS = {"fast_waves": 4, "eri_queue": "none", "chunk_mib": 16384, "cholesky_block": 32, "wave_max_l": 7,
            "mfma_min_l": 9, "mfma_variant": 0, "small_max_l": 2}
# knobs that do not change any sum: candidates must match bitwise
EXACT = ("eri_queue", "chunk_mib", "wave_max_l", "mfma_variant")


@dataclass
class Trial:
    knob: str
    value: object
    seconds: float  # best of the timed repetitions
    max_abs_diff: float  # against the default configuration on the check sample
    accepted: bool


@dataclass
class Choice:
    workload: str
    values: dict = field(default_factory=dict)
    trials: list = field(default_factory=list)

    def as_dict(self) -> dict:
        return {"workload": self.workload, "values": dict(self.values), "trials": [asdict(t) for t in self.trials]}


def apply(ctx: "_native.Context", values: dict) -> None:
    """Set the tuned values (any subset of the knobs) on a context."""
    if "fast_waves" in values:
        ctx.set_fast_waves(int(values["fast_waves"]))
    if "eri_queue" in values:
        ctx.set_eri_queue(str(values["eri_queue"]))
    if "chunk_mib" in values:
        ctx.set_chunk_mib(int(values["chunk_mib"]))
    if "cholesky_block" in values:
        ctx.set_cholesky_block(int(values["cholesky_block"]))
    if "wave_max_l" in values:
        ctx.set_wave_max_l(int(values["wave_max_l"]))
    if "mfma_min_l" in values:
        ctx.set_mfma_min_l(int(values["mfma_min_l"]))
    if "mfma_variant" in values:
        ctx.set_mfma_variant(int(values["mfma_variant"]))
    if "small_max_l" in values:
        ctx.set_small_max_l(int(values["small_max_l"]))


def _best_of(fn, reps: int) -> float:
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t)
    return best


def tune_eri(basis: "_native.DeviceBasis", quartets, field6, knobs=("fast_waves", "eri_queue"), reps: int = 2,
             check_quartets: int = 2000, tolerance: float = 1e-11, workload: str = "") -> Choice:
    """Tune the ERI knobs on `quartets` (outputs kept on the device while timing).

    Coordinate descent: each knob in turn takes the fastest accepted value
    while the others keep their current choice. Each candidate gets one
    untimed call, then the best of `reps` timed calls.
    """
    ctx = basis.ctx
    q = np.ascontiguousarray(np.asarray(quartets, dtype=np.uint32).reshape(-1, 4))
    total = int(basis.quartet_sizes(q).sum())
    sample = q[: max(1, min(check_quartets, len(q)))]
    values = {k: DEFAULTS[k] for k in knobs}
    apply(ctx, values)
    reference = basis.eri_batch(sample, field6)
    choice = Choice(workload=workload or f"{len(q)} quartets")
    out = ctx.alloc(max(total, 1) * 16)
    try:
        for knob in knobs:
            space = {"fast_waves": FAST_WAVES, "eri_queue": ERI_QUEUE, "chunk_mib": CHUNK_MIB,
                     "wave_max_l": WAVE_MAX_L, "mfma_min_l": MFMA_MIN_L, "mfma_variant": MFMA_VARIANT,
                     "small_max_l": SMALL_MAX_L}[knob]
            exact = knob in EXACT
            # an exact knob is compared with the integrals of the choices made so far
            base = basis.eri_batch(sample, field6) if exact else reference
            best = None
            for value in space:
                trial_values = dict(values, **{knob: value})
                apply(ctx, trial_values)
                got = basis.eri_batch(sample, field6)
                diff = float(np.max(np.abs(got - base))) if got.size else 0.0
                ok = diff == 0.0 if exact else diff <= tolerance
                secs = float("inf")
                if ok:
                    basis.eri_batch(q, field6, device_out=out)
                    ctx.synchronize()

                    def run():
                        basis.eri_batch(q, field6, device_out=out)
                        ctx.synchronize()

                    secs = _best_of(run, reps)
                choice.trials.append(Trial(knob, value, secs, diff, ok))
                if ok and (best is None or secs < best[1]):
                    best = (value, secs)
            if best is not None:
                values[knob] = best[0]
            apply(ctx, values)
    finally:
        ctx.free(out)
    choice.values = values
    return choice


def tune_cholesky(basis: "_native.DeviceBasis", field6, threshold: float, stop_after: int,
                  blocks=CHOLESKY_BLOCK, workload: str = "") -> Choice:
    """Tune the pivot block on a run cut at `stop_after` pivots.

    A candidate is accepted only if its pivots and factors are bitwise those
    of the default block size.
    """
    ctx = basis.ctx
    choice = Choice(workload=workload or f"Cholesky, {stop_after} pivots")

    def run(block):
        ctx.set_cholesky_block(int(block))
        t = time.perf_counter()
        ch = _native.cholesky_run(basis, field6, threshold, 0, stop_after)
        secs = time.perf_counter() - t
        piv = ch.pivots()
        last = ch.vectors(ch.rank - 1, 1) if ch.rank else np.zeros(0)
        ch.close()
        return secs, piv, last

    _, ref_piv, ref_last = run(DEFAULTS["cholesky_block"])
    best = None
    for block in blocks:
        secs, piv, last = run(block)
        ok = piv.shape == ref_piv.shape and np.array_equal(piv, ref_piv) and np.array_equal(last, ref_last)
        diff = 0.0 if ok else float("inf")
        if ok and last.size:
            diff = float(np.max(np.abs(last - ref_last)))
        choice.trials.append(Trial("cholesky_block", int(block), secs if ok else float("inf"), diff, ok))
        if ok and (best is None or secs < best[1]):
            best = (int(block), secs)
    value = best[0] if best else DEFAULTS["cholesky_block"]
    ctx.set_cholesky_block(value)
    choice.values = {"cholesky_block": value}
    return choice


def cache_path() -> Path:
    env = os.environ.get("GIAOINT_AMD_TUNE_CACHE")
    return Path(env) if env else Path.home() / ".cache" / "giaoint_amd" / "tuning.json"


def _key(device_name: str, workload: str) -> str:
    return f"{device_name}|{workload}"


def save(device_name: str, choice: Choice, path: Path | None = None) -> Path:
    p = Path(path) if path else cache_path()
    p.parent.mkdir(parents=True, exist_ok=True)
    data = json.loads(p.read_text()) if p.exists() else {}
    data[_key(device_name, choice.workload)] = choice.values
    p.write_text(json.dumps(data, indent=1, sort_keys=True))
    return p


def load(device_name: str, workload: str, path: Path | None = None) -> dict | None:
    p = Path(path) if path else cache_path()
    if not p.exists():
        return None
    return json.loads(p.read_text()).get(_key(device_name, workload))
