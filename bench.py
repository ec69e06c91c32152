"""Benchmark: shell quartets/s and FP64 GFLOP/s of the London ERI path on one MI355X.

One step = one evaluation of the whole synthetic batch of 200,000 (dd|dd)
Cartesian shell quartets (four private shells per quartet, contraction degree
3 => 81 primitive quartets each, complex128 output of 4.15 GB kept in HBM).
Inputs (shell tables) are resident on the device before timing; the output
is written to device memory. Prints one JSON line. With --cholesky-rank
the line also carries the streamed Cholesky of the 30-atom set (wall time,
rank, device time split).

The CPU baseline times GIAOINT's own OpenMP path compiled from its sources
(oracle/_ref, kind "reference"), or the C restatement when that library is
absent (kind "port"), on the first quartets of the same batch using every
host core of this process.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector / matrix (dense)
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# Per (dd|dd) K=3 quartet of fast_transform_kernel<2,2,2,2>, by workgroup
# size (4 waves: 7 + 2 bra pairs in two passes; the default cap gives 5 waves
# and one pass): HBM bytes from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs
# scaled by a copy-kernel calibration, and FP64 operations issued from the
# SQ_INSTS_VALU_{FMA,ADD,MUL,TRANS}_F64 counters (64 lanes, FMA = 2)
# (profiles/fast_transform_kernel_dddd_200k_pmc.md, profiles/fp64_mfma_lds_counters.md).
TRANSFORM_HBM_BYTES_PER_QUARTET = {4: 477_171, 5: 370_995}
TRANSFORM_FP64_ISSUED_PER_QUARTET = {4: 1.1159e7}
TRANSFORM_TRAFFIC_NOTE = ("HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE (x2.00) + WRITE_SIZE (x1.00), "
                          "factors from a 256 MiB copy; {b:,} B per (dd|dd) quartet at {w} waves x quartets per launch")
METRIC = "shell quartets/s + FP64 GFLOP/s vs peak; Cholesky wall time, 1 MI355X"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--quartets", type=int, default=200_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain kernel")
    ap.add_argument("--chunk-mib", type=int, default=8192,
                    help="batched-path chunk size of the headline leg in MiB (0: autotuned with the workgroup size); "
                         "fixed by default so that the contraction launches of every run have one size")
    ap.add_argument("--cholesky-rank", type=int, default=0,
                    help="Cholesky leg: stop after this many pivots (0: run to convergence, -1: skip the leg)")
    ap.add_argument("--cholesky-block", type=int, default=0,
                    help="candidates per pivot block (default: autotuned, else the library's)")
    ap.add_argument("--fast-waves", type=int, default=-1,
                    help="workgroup waves of the class-specialized kernel for the headline batch "
                         "(0 = library default, 4..8; default: autotuned)")
    ap.add_argument("--pipeline", choices=("overlap", "serial"), default="overlap",
                    help="producer kernels of the batched path beside the contraction (default) or serially")
    ap.add_argument("--no-host-output", action="store_true",
                    help="skip the PCIe-inclusive run into host memory (profiles of the timed steps only)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch parameters instead of autotuning them")
    ap.add_argument("--tune-cholesky-rank", type=int, default=1500,
                    help="pivots of each Cholesky run the autotuner times")
    ap.add_argument("--high-l-gggg", type=int, default=50_000, help="(gg|gg) quartets of the high-l leg (0 skips)")
    ap.add_argument("--high-l-ffdd", type=int, default=200_000, help="(ff|dd) quartets of the high-l leg (0 skips)")
    ap.add_argument("--mixed", type=int, default=2_000_000,
                    help="mixed-class leg: quartets of the l<=3, degree 1..8 workload (0 skips it)")
    ap.add_argument("--no-derivatives", action="store_true", help="skip the ERI centre-derivative leg")
    ap.add_argument("--no-cpu-legs", action="store_true",
                    help="skip the CPU baselines of the other workloads (water tensor, mixed, high l, Cholesky)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def physical_cores() -> int:
    """Physical cores available to this process, counted as the reference
    counts them for its default thread budget (python/giao_integrals/
    _threads.py:33-64): distinct (package, core) pairs over the CPUs of the
    affinity mask; 1 when the topology cannot be read."""
    try:
        cores = set()
        for cpu in os.sched_getaffinity(0):
            topo = Path(f"/sys/devices/system/cpu/cpu{cpu}/topology")
            cores.add(((topo / "physical_package_id").read_text().strip(), (topo / "core_id").read_text().strip()))
        return max(1, len(cores))
    except (AttributeError, OSError):
        return 1


def _checker():
    from oracle import pyoracle

    kind = "reference" if pyoracle.available("reference") else "port"
    return kind, pyoracle.load(kind)


def _time_quartets(chk, batch, rows, threads: int) -> float:
    """Seconds the checker takes for the quartets `rows` of a batch (each
    quartet has four private shells), one OpenMP call."""
    from oracle import pyoracle

    rows = np.asarray(rows, dtype=np.int64)
    sub = batch.shells.subset(np.concatenate([batch.quartets[i] for i in rows]).astype(np.int64))
    basis = pyoracle.make_basis(sub.shell_list(), chk)
    q = np.arange(4 * len(rows), dtype=np.uint32).reshape(-1, 4)
    f = chk.field6(batch.B, batch.origin)
    t = time.perf_counter()
    chk.eri_batch(basis, q, f, threads=threads)
    return time.perf_counter() - t


def _sized_sample(chk, batch, rows, threads: int, target_s: float):
    """Times a calibration prefix of `rows`, then the longest prefix expected
    to take about target_s; returns (rows timed, seconds)."""
    n0 = min(len(rows), max(threads, 16))
    dt0 = _time_quartets(chk, batch, rows[:n0], threads)
    n = int(min(len(rows), max(n0, target_s / max(dt0 / n0, 1e-9))))
    n = max(min(n0, n), (n // threads) * threads)
    return n, _time_quartets(chk, batch, rows[:n], threads)


def cpu_baseline(batch, target_s: float):
    """The headline (dd|dd) batch on the host cores: the reference's own
    OpenMP path (oracle/_ref, kind "reference"), or the C restatement when
    that library is absent (kind "port"), on a prefix of the same batch."""
    kind, chk = _checker()
    threads = physical_cores()
    n, dt = _sized_sample(chk, batch, np.arange(len(batch.quartets)), threads, target_s)
    return {
        "value": n / dt,
        "unit": "shell quartets/s",
        "cores": threads,
        "kind": kind,
        "sample": f"first {n} of the same (dd|dd) batch, OpenMP schedule(dynamic,1) on {threads} physical cores, "
                  f"{dt:.1f} s",
    }


def cpu_baselines(mixed_batch, target_s: float) -> dict:
    """The other workloads on the host cores, each on a bounded sample
    (rates are per sample, labelled as such):

    * water cc-pVDZ, full 25^4 tensor at B = (0,0,0.1): ms per tensor;
    * mixed classes: a stratified sample (every k-th quartet of the 2M batch);
    * (gg|gg): one quartet on one core (about 40 s of the reference's CPU
      time each); (ff|dd): a prefix of the 200k batch on every core;
    * 30-atom Cholesky: the diagonal and the first 32 pivots of the checker
      driver (oracle/cholesky_driver.py, the reference's driver restated, over
      the checker's integrals), seconds per pivot, and the full run's time
      extrapolated from it (labelled as an extrapolation)."""
    from giaoint_amd import workloads
    from oracle import cholesky_driver, pyoracle

    kind, chk = _checker()
    threads = physical_cores()
    out = {}
    # water cc-pVDZ full tensor
    sh, B, O = workloads.water_ccpvdz()
    basis = pyoracle.make_basis(sh.shell_list(), chk)
    f = chk.field6(B, O)
    reps, t = 0, time.perf_counter()
    while reps < 3 or (time.perf_counter() - t < 2.0 and reps < 50):
        chk.eri_tensor(basis, f, threads)
        reps += 1
    dt = (time.perf_counter() - t) / reps
    out["water_ccpvdz_tensor"] = {"value": dt * 1e3, "unit": "ms per 25^4 tensor", "cores": threads, "kind": kind,
                                  "sample": f"{reps} full tensors (637 canonical quartets each)"}
    log(f"cpu: water tensor {dt * 1e3:.1f} ms")
    # mixed classes: stratified sample
    if mixed_batch is not None:
        stride = max(1, len(mixed_batch.quartets) // 4000)
        rows = np.arange(0, len(mixed_batch.quartets), stride)
        n, dt = _sized_sample(chk, mixed_batch, rows, threads, target_s)
        out["mixed_class"] = {"value": n / dt, "unit": "shell quartets/s", "cores": threads, "kind": kind,
                              "sample": f"{n} quartets, every {stride}th of the {len(mixed_batch.quartets)}-quartet "
                                        f"mixed batch, {dt:.1f} s"}
        log(f"cpu: mixed {n / dt:.1f} quartets/s")
    # high l
    g, fd = workloads.precision_batches(n_gggg=1, n_ffdd=2000)
    dt = _time_quartets(chk, g, [0], 1)
    out["gggg"] = {"value": 1.0 / dt, "unit": "shell quartets/s", "cores": 1, "kind": kind,
                   "sample": f"1 (gg|gg) K=2^4 quartet on one core, {dt:.1f} s"}
    log(f"cpu: gggg {dt:.1f} s per quartet")
    n, dt = _sized_sample(chk, fd, np.arange(len(fd.quartets)), threads, target_s / 2)
    out["ffdd"] = {"value": n / dt, "unit": "shell quartets/s", "cores": threads, "kind": kind,
                   "sample": f"first {n} (ff|dd) K=2^4 quartets, {dt:.1f} s"}
    # 30-atom Cholesky prefix
    sh, B, O = workloads.cholesky_atoms()
    basis = pyoracle.make_basis(sh.shell_list(), chk)
    f = chk.field6(B, O)
    desc = [((l + 1) * (l + 2) // 2, 1, 1) for l in basis.L]
    t = time.perf_counter()
    chk.cholesky_diagonal(basis, [(a, c) for a in range(len(desc)) for c in range(a, len(desc))], f, threads)
    diag_s = time.perf_counter() - t
    t = time.perf_counter()
    _, piv, _, _ = cholesky_driver.cholesky_eri(
        basis.ao_offsets, desc, lambda pairs: chk.cholesky_diagonal(basis, pairs, f, threads),
        lambda pv: chk.cholesky_columns(basis, pv, f, threads), real=False, single_center=False, threshold=1e-8,
        max_pivots=32)
    run_s = time.perf_counter() - t
    per_pivot = (run_s - diag_s) / len(piv)
    out["cholesky_30_atoms"] = {
        "value": per_pivot, "unit": "s per pivot", "cores": threads, "kind": kind,
        "sample": f"diagonal ({diag_s:.2f} s) and the first {len(piv)} pivots of the checker driver over the "
                  f"checker's integrals, {run_s:.1f} s",
        "extrapolated_full_run_s": diag_s + per_pivot * 6348,
        "extrapolation_note": "per-pivot time of the first 32 pivots x rank 6348; the projection term grows with "
                              "the rank, so the full run would take longer",
    }
    log(f"cpu: cholesky {per_pivot:.3f} s per pivot")
    return out


def mixed_leg(n: int, tune: bool = True, tune_n: int = 500_000):
    """Mixed classes l <= 3 (weights (1 + sum l)^-2), contraction degree U{1..8}:
    the 2M-quartet workload. With `tune`, two kernel choices are autotuned on
    the first tune_n quartets: the largest total angular momentum whose
    generic-transform items run on one-wavefront workgroups, and the smallest
    whose classes run on the FP64 matrix cores; then one untimed and one
    timed call over the whole batch with the batched variant (class-specialized
    kernels where they exist, matrix cores from the tuned momentum, the
    generic kernel for the rest) and with the plain variant (the generic
    kernel for every class), inputs and output resident on the device.
    Returns the leg's record and the batch."""
    from giaoint_amd import _native, autotune, workloads

    t0 = time.perf_counter()
    batch = workloads.mixed_batch(n)
    log(f"mixed: {n} quartets generated in {time.perf_counter() - t0:.1f} s")
    sh = batch.shells
    coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
    ctx = _native.Context(0)
    db = _native.DeviceBasis(ctx, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)
    total = int(db.quartet_sizes(batch.quartets).sum())
    f = _native.field6(batch.B, batch.origin)
    flops = workloads.batch_flops(sh, batch.quartets)
    tuned = None
    if tune:
        m = min(n, tune_n)
        tuned = autotune.tune_eri(db, batch.quartets[:m], f, knobs=("fast_waves", "wave_max_l", "mfma_min_l"), reps=1,
                                  workload=f"mixed l<=3 degree 1..8, first {m} of {n} quartets")
        log(f"mixed: autotuned {tuned.values} " +
            " ".join(f"{t.value}={t.seconds:.3f}s" for t in tuned.trials))
    out = ctx.alloc(total * 16)
    timings = {}
    try:
        for variant in ("batched", "plain"):
            ctx.set_variant(variant)
            # one untimed call over the whole batch: the first call at a new
            # size grows the context's device and page-locked buffers
            db.eri_batch(batch.quartets, f, device_out=out)
            ctx.synchronize()
            t = time.perf_counter()
            db.eri_batch(batch.quartets, f, device_out=out)
            ctx.synchronize()
            timings[variant] = (time.perf_counter() - t, ctx.stats())
            log(f"mixed ({variant}): {n} quartets in {timings[variant][0]:.2f} s")
    finally:
        ctx.set_variant("batched")
        ctx.free(out)
    dt, st = timings["batched"]
    extra = {"autotune": tuned.as_dict()} if tuned else {}
    return {**extra,
        "workload": f"{n} quartets, classes l<=3 weighted (1+sum l)^-2, degree U{{1..8}}, seed 3, "
                    f"{st['primitive_quartets']} primitive quartets",
        "quartets_per_s": n / dt,
        "seconds": dt,
        "tflops_convention": flops / dt / 1e12,
        "launches": st["launches"],
        "plain_kernel": {"quartets_per_s": n / timings["plain"][0], "seconds": timings["plain"][0],
                         "tflops_convention": flops / timings["plain"][0] / 1e12},
        "speedup_batched_vs_plain": timings["plain"][0] / dt,
    }, batch


def high_l_leg(n_gggg: int, n_ffdd: int, probe_values: int = 4_000_000, tune: bool = True):
    """The high-angular-momentum precision workloads: (gg|gg) and (ff|dd),
    contraction degree 2, exponents 10^U(log10 0.05, log10 5e3), each with
    its E.R.E contraction on FP64 VALU (the generic kernel for (gg|gg), the
    class-specialized (dd|ff) kernel for (ff|dd)), on the FP64 matrix
    cores (v_mfma_f64_16x16x4f64, dense complex Hermite products) and in
    split FP32 on the FP32 matrix cores (hi/lo products on
    v_mfma_f32_16x16x4f32, accumulated in FP32 per step of four and
    compensated into FP64). A warmup call, then one timed call per variant,
    output resident on the device.
    Errors: max and RMS of |variant - FP64 VALU| over the first
    `probe_values` integrals of the batch (the VALU path is held to the C
    oracle at 1e-11 by the GPU tests)."""
    from giaoint_amd import _native, workloads

    g, fd = workloads.precision_batches(n_gggg=n_gggg, n_ffdd=n_ffdd)
    out = {}
    mfma_var = 0  # compile-time variant of the matrix-core kernels, tuned on the first (gg|gg) quartets
    for name, batch in (("gggg", g), ("ffdd", fd)):
        n = len(batch.quartets)
        if n == 0:
            continue
        sh = batch.shells
        coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
        ctx = _native.Context(0)
        db = _native.DeviceBasis(ctx, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)
        total = int(db.quartet_sizes(batch.quartets).sum())
        dev = ctx.alloc(total * 16)
        f = _native.field6(batch.B, batch.origin)
        flops = workloads.batch_flops(sh, batch.quartets)
        probe = min(total, probe_values)
        rec = {"workload": batch.description, "quartets": n, "variants": {}}
        ref = None
        if tune and name == "gggg":
            from giaoint_amd import autotune

            ctx.set_mfma("all")
            m = min(n, 4000)
            choice = autotune.tune_eri(db, batch.quartets[:m], f, knobs=("mfma_variant",), reps=2, check_quartets=16,
                                       workload=f"(gg|gg) K=2, first {m} quartets, FP64 matrix cores")
            mfma_var = int(choice.values["mfma_variant"])
            rec["autotune"] = choice.as_dict()
            log(f"gggg: MFMA variant {mfma_var} " + " ".join(f"{t.value}={t.seconds:.3f}s" for t in choice.trials))
        ctx.set_mfma_variant(mfma_var)
        try:
            for variant, mode in (("fp64_valu", "off"), ("fp64_mfma", "all"), ("split_fp32_mfma", "split32")):
                ctx.set_mfma(mode)
                db.eri_batch(batch.quartets[: min(n, 256)], f, device_out=dev)
                ctx.synchronize()
                t = time.perf_counter()
                db.eri_batch(batch.quartets, f, device_out=dev)
                ctx.synchronize()
                dt = time.perf_counter() - t
                vals = ctx.to_host(dev, np.empty(probe, dtype=np.complex128))
                v = {"quartets_per_s": n / dt, "seconds": dt, "tflops_convention": flops / dt / 1e12}
                if ref is None:
                    ref = vals
                else:
                    err = np.abs(vals - ref)
                    v["max_abs_err_vs_fp64_valu"] = float(err.max())
                    v["rms_err_vs_fp64_valu"] = float(np.sqrt(np.mean(err**2)))
                    v["max_rel_err_vs_fp64_valu"] = float((err / np.maximum(np.abs(ref), 1e-300)).max())
                    v["within_1e-10"] = bool(err.max() <= 1e-10)
                    v["max_abs_value"] = float(np.abs(ref).max())
                    v["error_sample"] = f"first {probe} integrals ({probe * n // max(total, 1)} quartets)"
                rec["variants"][variant] = v
                log(f"{name} {variant}: {n} quartets in {dt:.3f} s")
        finally:
            ctx.set_mfma("default")
            ctx.free(dev)
        # the headline rate of the leg is the fastest variant within the 1e-10 bar
        ok = {k: v for k, v in rec["variants"].items() if v.get("within_1e-10", True)}
        best = max(ok.items(), key=lambda kv: kv[1]["quartets_per_s"])
        rec["quartets_per_s"] = best[1]["quartets_per_s"]
        rec["fastest"] = best[0]
        out[name] = rec
    return out


def derivative_leg(target_s: float, cpu: bool):
    """ERI centre derivatives (compute_eri_derivative_blocks) of every
    canonical shell quartet of water in cc-pVTZ (65 AOs, general
    contractions, up to f) at B = (0.05, -0.03, 0.2), gauge (0.37, -0.29,
    0.43), blocks resident on the device: a warmup call, then one timed call.
    CPU: the reference's own derivatives.cpp (oracle/_ref) over every k-th
    quartet, sized to about target_s on the physical cores."""
    from giaoint_amd import core, workloads
    from giaoint_amd.eri import _quartets
    from giaoint_amd.molecule import Molecule

    atoms = [("O", (0.0, 0.1, 0.2)), ("H", (0.1, 1.4, -0.9)), ("H", (-0.2, -1.3, -1.0))]
    field = core.MagneticField((0.05, -0.03, 0.2), (0.37, -0.29, 0.43))
    mol = Molecule(atoms, workloads.DATA / "cc-pvtz-h-o.json", unit="bohr", field=field)
    dev = mol.basis._dev()
    ctx = dev.ctx
    qs = np.array(list(_quartets(len(mol.basis.shells), None)), dtype=np.uint32)
    total = 12 * int(dev.quartet_sizes(qs).sum())
    f = np.asarray(field._field6(), dtype=np.float64)
    out = ctx.alloc(total * 16)
    try:
        dev.eri_derivative_batch(qs, f, device_out=out)
        ctx.synchronize()
        t = time.perf_counter()
        dev.eri_derivative_batch(qs, f, device_out=out)
        ctx.synchronize()
        dt = time.perf_counter() - t
        device_ms = ctx.stats()["total_ms"]
        checksum = float(np.abs(ctx.to_host(out, np.empty(min(total, 1 << 22), dtype=np.complex128))).sum())
    finally:
        ctx.free(out)
    rec = {"workload": f"water cc-pVTZ ({mol.ao_count} AOs, {len(mol.basis.shells)} shells), all {len(qs)} "
                       "canonical shell quartets, 12 centre derivatives each, London phase at B = (0.05, -0.03, 0.2)",
           "quartets_per_s": len(qs) / dt, "seconds": dt, "device_ms": device_ms,
           "output_gb_per_s": total * 16 / dt / 1e9, "output_values": total, "checksum": checksum}
    log(f"derivatives: {len(qs)} quartets in {dt:.3f} s")
    if cpu:
        try:
            from oracle import pyoracle

            kind, chk = _checker()
            threads = physical_cores()
            # the molecule's shells with their contraction-normalized coefficients
            shells = [(s.center, s.angular_momentum, s._exps, s._coefs) for s in mol.basis.shells]
            arrays = pyoracle.BasisArrays(shells)
            t0 = time.perf_counter()
            chk.eri_derivative_blocks(arrays, qs[::97][:threads], f, threads=threads)
            dt0 = time.perf_counter() - t0
            per = dt0 / min(threads, len(qs[::97]))
            step = max(1, int(np.ceil(len(qs) * per / target_s)))
            sample = qs[::step]
            t0 = time.perf_counter()
            chk.eri_derivative_blocks(arrays, sample, f, threads=threads)
            cdt = time.perf_counter() - t0
            rec["cpu_baseline"] = {"value": len(sample) / cdt, "unit": "derivative shell quartets/s", "cores": threads,
                                   "kind": kind, "sample": f"every {step}th canonical quartet ({len(sample)}), "
                                                           f"{cdt:.1f} s"}
            rec["gpu_over_cpu"] = rec["quartets_per_s"] / rec["cpu_baseline"]["value"]
        except Exception as exc:  # the baseline is reported, never required
            rec["cpu_baseline"] = {"error": str(exc)}
    return rec


def projection_roofline(nbytes: float, flops: float, block_ms: float, unblocked_bytes: float) -> dict:
    """Roofline of the Cholesky block projections: factor bytes streamed from
    HBM and complex multiply-adds (8 FLOP each, one per candidate projected
    per factor element); the bound is whichever takes longer at peak."""
    sec = block_ms * 1e-3
    t_hbm = nbytes / (HBM_PEAK_GBS * 1e9)
    t_fp64 = flops / (FP64_PEAK_TFLOPS * 1e12)
    common = {
        "bytes": nbytes,
        "flops": flops,
        "hbm_gbs": nbytes / sec / 1e9,
        "fp64_tflops": flops / sec / 1e12,
        "block_ms": block_ms,
        "note": "device time of candidate selection + projection; FLOPs count the candidates projected x factor "
                f"rows x n^2 complex MACs; one GEMV per pivot would read {unblocked_bytes / nbytes:.1f}x as many bytes",
    }
    if t_fp64 >= t_hbm:
        return {"bound": "fp64", "achieved": flops / sec / 1e12, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": t_fp64 / sec, **common}
    return {"bound": "hbm", "achieved": nbytes / sec / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": t_hbm / sec, **common}


def cholesky_leg(rank_cap: int, block: int = 0, tune_rank: int = 0, graphs_rank: int = 1500):
    """Streamed pivoted Cholesky of the 30-atom 4s3p2d1f set (1050 AOs), tau = 1e-8.

    First the column launches as hipGraphs against direct launches on the
    first graphs_rank pivots (wall time of each, identical pivots required);
    then the device driver in steps of 1000 pivots (a progress line after
    each) until convergence, or until rank_cap pivots when rank_cap > 0.
    """
    from giaoint_amd import _native, workloads

    sh, B, O = workloads.cholesky_atoms()
    coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
    ctx = _native.Context(0)
    if block:
        ctx.set_cholesky_block(block)
    db = _native.DeviceBasis(ctx, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)
    field = _native.field6(B, O)
    tuned = None
    if not block and tune_rank > 0:
        from giaoint_amd import autotune

        tuned = autotune.tune_cholesky(db, field, 1e-8, tune_rank,
                                       workload=f"30-atom 4s3p2d1f Cholesky, first {tune_rank} pivots")
        log(f"cholesky: autotuned {tuned.values} " +
            " ".join(f"B={t.value}:{t.seconds:.2f}s" for t in tuned.trials))
    graphs_cmp = None
    if graphs_rank > 0:
        # the first graphs_rank pivots with the column launches as hipGraphs and direct
        graphs_cmp = {"prefix_pivots": graphs_rank}
        piv = {}
        for on in (False, True):
            ctx.set_cholesky_graphs(on)
            t = time.perf_counter()
            pre = _native.cholesky_run(db, field, 1e-8, 0, graphs_rank)
            graphs_cmp["prefix_graphs_s" if on else "prefix_direct_s"] = time.perf_counter() - t
            piv[on] = pre.pivots()
            pre.close()
        graphs_cmp["prefix_pivots_identical"] = bool(np.array_equal(piv[True], piv[False]))
        ctx.set_cholesky_graphs(False)
        log(f"cholesky: first {graphs_rank} pivots, direct {graphs_cmp['prefix_direct_s']:.2f} s, "
            f"graphs {graphs_cmp['prefix_graphs_s']:.2f} s")
    step = 1000
    t0 = time.perf_counter()
    first = step if rank_cap <= 0 else min(step, rank_cap)
    ch = _native.cholesky_run(db, field, 1e-8, 0, first)
    while not ch.converged and (rank_cap <= 0 or ch.rank < rank_cap):
        log(f"cholesky: rank {ch.rank} max residual {ch.max_residual:.3e} {time.perf_counter() - t0:.1f} s")
        ch.resume(ch.rank + step if rank_cap <= 0 else min(ch.rank + step, rank_cap))
    wall = time.perf_counter() - t0
    n = db.nao
    m = ch.rank
    t = ch.timings_ms
    t_col = t["column"]
    # a device-resident consumer of the factors: J[mu,nu] = sum (mu nu|lam sig) D[lam,sig]
    rng = np.random.default_rng(0)
    dens = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    tj = time.perf_counter()
    ch.coulomb(dens)
    coulomb_s = time.perf_counter() - tj
    tk = time.perf_counter()
    ch.exchange(dens)
    exchange_s = time.perf_counter() - tk
    exchange_ms = ctx.stats()["transform_ms"]
    # projection: each pivot block reads the factors computed before it once
    blk = ch.blocks
    proj_bytes = 16.0 * n * n * blk["factor_rows_read"]
    proj_flops = 8.0 * n * n * blk["candidate_rows"]
    unblocked_bytes = 16.0 * n * n * m * (m - 1) / 2
    result = {
        **({"autotune": tuned.as_dict()} if tuned else {}),
        "workload": "30 atoms, min spacing 1.8 bohr in a 14-bohr box (seed 5), uncontracted 4s3p2d1f "
                    f"(alpha_i = 0.05*3^i), {n} Cartesian AOs, B=(0.05,0,0.1) a.u., tau=1e-8, complex128 factors",
        "rank": m,
        "converged": ch.converged,
        "max_residual": ch.max_residual,
        "tolerance": ch.tolerance,
        "wall_s": wall,
        "device_ms": t,
        "coulomb_contraction": {
            "seconds": coulomb_s,
            "factor_passes_gbs": 2 * 16.0 * n * n * m / coulomb_s / 1e9,
            "note": "J = sum_P L_P (sum conj(L_P^T) D): two passes over the device factors, density and J over PCIe",
        },
        "exchange_contraction": {
            "seconds": exchange_s,
            "device_ms": exchange_ms,
            "tflops": 2 * 8.0 * n**3 * m / (exchange_ms * 1e-3) / 1e12,
            "note": "K = sum_P L_P D L_P^H: two rocBLAS zgemm of order n per factor (8 FLOP per complex MAC), "
                    "density and K over PCIe",
        },
        "column_graphs": graphs_cmp,
        "pivot_blocks": blk["blocks"],
        "pivots_per_block": m / max(blk["blocks"], 1),
        "candidates_projected_per_block": blk["projected"] / max(blk["blocks"], 1),
        "projection": projection_roofline(proj_bytes, proj_flops, blk["block_ms"], unblocked_bytes),
    }
    direct_pivots = ch.pivots()
    ch.close()
    if graphs_cmp is not None and rank_cap <= 0:
        # the whole run once more with the column launches as hipGraphs (the
        # direct run's factors released first: two sets of 30-atom factors do
        # not fit on the device beside each other)
        ctx.set_cholesky_graphs(True)
        t = time.perf_counter()
        chg = _native.cholesky_run(db, field, 1e-8)
        graphs_cmp["full_graphs_s"] = time.perf_counter() - t
        graphs_cmp["full_direct_s"] = wall
        graphs_cmp["full_pivots_identical"] = bool(np.array_equal(chg.pivots(), direct_pivots))
        graphs_cmp["full_graph_use"] = chg.graphs
        graphs_cmp["full_column_ms"] = {"graphs": chg.timings_ms["column"], "direct": t_col}
        chg.close()
        ctx.set_cholesky_graphs(False)
        log(f"cholesky: whole run with graphs {graphs_cmp['full_graphs_s']:.2f} s (direct {wall:.2f} s)")
    return result


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is designed but not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    from giaoint_amd import _native, workloads

    t0 = time.perf_counter()
    batch = workloads.dddd_batch(args.quartets)
    sh = batch.shells
    coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
    log(f"workload generated in {time.perf_counter() - t0:.1f} s; field B={batch.B}")
    ctx = _native.Context(0)
    if args.chunk_mib:
        ctx.set_chunk_mib(args.chunk_mib)
    dbasis = _native.DeviceBasis(ctx, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)
    n = len(batch.quartets)
    total = int(dbasis.quartet_sizes(batch.quartets).sum())
    out = ctx.alloc(total * 16)
    field = _native.field6(batch.B, batch.origin)
    flops = workloads.batch_flops(sh, batch.quartets)
    contract_flops = workloads.batch_flops(sh, batch.quartets, part="contract")
    executed = workloads.batch_executed_flops(sh, batch.quartets)

    def step():
        dbasis.eri_batch(batch.quartets, field, device_out=out)
        return ctx.stats()

    tuned = None
    ctx.set_pipeline(args.pipeline)
    if args.fast_waves >= 0:
        ctx.set_fast_waves(args.fast_waves)
    elif not args.no_autotune:
        from giaoint_amd import autotune

        ctx.free(out)  # the tuner holds its own output buffer
        knobs = ("fast_waves",) if args.chunk_mib else ("fast_waves", "chunk_mib")
        tuned = autotune.tune_eri(dbasis, batch.quartets, field, knobs=knobs,
                                  workload=f"(dd|dd) K=3, {n} quartets")
        log(f"autotuned {tuned.values}: " + " ".join(f"{t.knob}={t.value}:{t.seconds:.3f}s" for t in tuned.trials))
        out = ctx.alloc(total * 16)

    # plain FP64 kernel (generic, one launch per class) on the same batch, for comparison
    plain = None
    if not args.no_plain:
        ctx.set_variant("plain")
        step()
        plain = step()
        ctx.set_variant("batched")
        log(f"plain kernel: transform {plain['transform_ms']:.1f} ms seed {plain['seed_ms']:.1f} ms "
            f"total {plain['total_ms']:.1f} ms")
    for i in range(args.warmup):
        s = step()
        log(f"warmup {i}: transform {s['transform_ms']:.1f} ms seed {s['seed_ms']:.1f} ms")
    ctx.synchronize()
    stats = []
    t = time.perf_counter()
    for i in range(args.steps):
        stats.append(step())
        log(f"step {i}: transform {stats[-1]['transform_ms']:.1f} ms seed {stats[-1]['seed_ms']:.1f} ms "
            f"device {stats[-1]['total_ms']:.1f} ms call {stats[-1]['host_ms']:.1f} ms")
    ctx.synchronize()
    wall = time.perf_counter() - t
    ms_step = 1e3 * wall / args.steps
    transform_ms = float(np.mean([s["transform_ms"] for s in stats]))
    seed_ms = float(np.mean([s["seed_ms"] for s in stats]))
    produce_ms = float(np.mean([s["produce_ms"] for s in stats]))
    contract_ms = float(np.mean([s["contract_ms"] for s in stats]))
    contract_launches = stats[-1]["contract_launches"]
    launches = stats[-1]["launches"]
    # the same batch delivered into host memory (PCIe-inclusive rate, not the headline value)
    host_rate = None
    if not args.no_host_output:
        host_out = np.empty(total, dtype=np.complex128)
        host_out.fill(0.0)  # fault the pages in before timing
        t = time.perf_counter()
        dbasis.eri_batch(batch.quartets, field, out=host_out)
        host_rate = n / (time.perf_counter() - t)
        del host_out
    # checksum of the resident output for run-to-run comparison
    probe = np.empty(min(total, 1 << 20), dtype=np.complex128)
    ctx.to_host(out, probe)
    # dominant kernel: the E.R.E contraction, timed with events on the stream it runs on
    avg_launch_ms = contract_ms / contract_launches
    # the contraction's own work: 8 FLOP per complex MAC over the nonzero
    # Hermite products (the executed count without the R0 part, which runs in
    # r0_kernel), a count the kernel cannot exceed at the FP64 peak
    contract_useful = executed - (flops - contract_flops)
    achieved = contract_useful / contract_launches / (avg_launch_ms * 1e-3) / 1e12
    convention_rate = contract_flops / contract_launches / (avg_launch_ms * 1e-3) / 1e12
    fast_waves = args.fast_waves if args.fast_waves >= 0 else (tuned.values["fast_waves"] if tuned else 0)
    waves_used = 4 if fast_waves == 4 else 5  # 9 bra pairs x 35 Hermite rows (workload dddd K=3)
    traffic = None
    if waves_used in TRANSFORM_HBM_BYTES_PER_QUARTET:
        traffic = TRANSFORM_HBM_BYTES_PER_QUARTET[waves_used] * n / contract_launches
    issued = None
    if waves_used in TRANSFORM_FP64_ISSUED_PER_QUARTET:
        issued = TRANSFORM_FP64_ISSUED_PER_QUARTET[waves_used] * n / contract_launches / (avg_launch_ms * 1e-3) / 1e12
    result = {
        "metric": METRIC,
        "value": n / (ms_step * 1e-3),
        "unit": "shell quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{n} (dd|dd) Cartesian shell quartets, contraction degree 3 (81 primitive quartets each), "
                        "centres U([0,6]^3) bohr, exponents 10^U(-1,4), coefficients N(0,1), "
                        f"B={tuple(round(x, 6) for x in batch.B)} a.u. (uniform in |B|<=0.5, seed 2), complex128 output",
            "quartets": n,
            "output_bytes": total * 16,
            "algorithmic_flops": flops,
            "kernel": "batched FP64 (class-specialized transform, wave-uniform Hermite products)",
        },
        "plain_kernel": None if plain is None else {
            "quartets_per_s": n / (plain["host_ms"] * 1e-3),
            "transform_ms": plain["transform_ms"],
            "seed_ms": plain["seed_ms"],
            "speedup_batched_vs_plain": plain["host_ms"] / float(np.mean([s["host_ms"] for s in stats])),
        },
        "host_output": {"quartets_per_s": host_rate,
                        "note": "one call writing the packed blocks to pageable host memory (PCIe included)"},
        "gflops": flops / (ms_step * 1e-3) / 1e9,
        "kernel_ms": {"seed": seed_ms, "pair_and_r0": produce_ms, "contract": contract_ms,
                      "launches_per_step": launches, "contract_launches_per_step": contract_launches},
        "roofline": {
            "bound": "fp64",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel": "gh::fast::fast_transform_kernel<2,2,2,2> (E.R.E contraction, FP64 VALU v_fmac_f64)",
            "avg_launch_ms": avg_launch_ms,
            "flops_per_launch": contract_useful / contract_launches,
            "flop_count": "8 per complex MAC over the nonzero complex Hermite products: every bra row (primitive "
                          "pair, Hermite index) meets every nonzero ket product of every ket pair, then every "
                          f"nonzero bra product meets every ket component ({contract_useful / n:.4g} per quartet); "
                          "the kernel must execute these, so achieved <= peak",
            "convention_tflops": convention_rate,
            "convention_note": "the dense-contraction convention (8 per complex MAC of the cheaper dense "
                               f"Hermite-to-Cartesian contraction, {contract_flops / n:.4g} per quartet) counts "
                               "structural zeros of the Hermite products that the kernel skips",
            "whole_path": {
                "executed_tflops": executed / (ms_step * 1e-3) / 1e12,
                "convention_tflops": flops / (ms_step * 1e-3) / 1e12,
                "note": "per step: the contraction's MACs plus the R0 recurrences (12 FLOP per entry, "
                        f"{(flops - contract_flops) / n:.4g} per quartet) over the step time",
            },
            "fp64_issued_tflops": issued,
            "fp64_issued_frac": None if issued is None else issued / FP64_PEAK_TFLOPS,
            "fp64_issued_note": "profile-derived constant: FP64 operations the kernel issues per quartet "
                                "(rocprofv3 SQ_INSTS_VALU_*_F64 counters, profiles/fp64_mfma_lds_counters.md) "
                                "over the live average launch time; includes padding of plan groups to 16 entries",
            "waves_per_workgroup": waves_used,
            "traffic_note": None if traffic is None else "profile-derived constant: " + TRANSFORM_TRAFFIC_NOTE.format(
                b=TRANSFORM_HBM_BYTES_PER_QUARTET[waves_used], w=waves_used),
        },
        "checksum": float(np.abs(probe).sum()),
    }
    if tuned is not None:
        result["autotune"] = tuned.as_dict()
    result["fast_waves"] = fast_waves
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(batch, args.cpu_seconds)
            result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
        except Exception as exc:  # the baseline is reported, never required
            result["cpu_baseline"] = {"error": str(exc)}
    ctx.free(out)
    del dbasis, ctx
    mixed = None
    if args.mixed > 0:
        result["mixed_class"], mixed = mixed_leg(args.mixed, tune=not args.no_autotune)
    if args.high_l_gggg > 0 or args.high_l_ffdd > 0:
        result["high_l"] = high_l_leg(args.high_l_gggg, args.high_l_ffdd, tune=not args.no_autotune)
    if not args.no_derivatives:
        result["eri_derivatives"] = derivative_leg(args.cpu_seconds, not args.no_cpu and not args.no_cpu_legs)
    if args.cholesky_rank >= 0:
        result["cholesky"] = cholesky_leg(args.cholesky_rank, args.cholesky_block,
                                          0 if args.no_autotune else args.tune_cholesky_rank)
    if not args.no_cpu and not args.no_cpu_legs:
        try:
            result["cpu_baselines"] = cpu_baselines(mixed, args.cpu_seconds)
            legs = result["cpu_baselines"]
            gpu_rates = {"mixed_class": result.get("mixed_class", {}).get("quartets_per_s"),
                         "gggg": result.get("high_l", {}).get("gggg", {}).get("quartets_per_s"),
                         "ffdd": result.get("high_l", {}).get("ffdd", {}).get("quartets_per_s")}
            for name, rate in gpu_rates.items():
                if rate and name in legs:
                    legs[name]["gpu_over_cpu"] = rate / legs[name]["value"]
        except Exception as exc:  # the baselines are reported, never required
            result["cpu_baselines"] = {"error": str(exc)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
