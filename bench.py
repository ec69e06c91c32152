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
TRANSFORM_HBM_BYTES_PER_QUARTET = {4: 466_622, 5: 370_995}
TRANSFORM_FP64_ISSUED_PER_QUARTET = {4: 1.4337e7}
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
    ap.add_argument("--chunk-mib", type=int, default=0, help="batched-path chunk size (default: library's)")
    ap.add_argument("--cholesky-rank", type=int, default=0,
                    help="Cholesky leg: stop after this many pivots (0: run to convergence, -1: skip the leg)")
    ap.add_argument("--cholesky-block", type=int, default=0,
                    help="candidates per pivot block (default: autotuned, else the library's)")
    ap.add_argument("--fast-waves", type=int, default=-1,
                    help="workgroup waves of the class-specialized kernel for the headline batch "
                         "(0 = library default, 4..8; default: autotuned)")
    ap.add_argument("--no-host-output", action="store_true",
                    help="skip the PCIe-inclusive run into host memory (profiles of the timed steps only)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch parameters instead of autotuning them")
    ap.add_argument("--tune-cholesky-rank", type=int, default=1500,
                    help="pivots of each Cholesky run the autotuner times")
    ap.add_argument("--high-l-gggg", type=int, default=50_000, help="(gg|gg) quartets of the high-l leg (0 skips)")
    ap.add_argument("--high-l-ffdd", type=int, default=200_000, help="(ff|dd) quartets of the high-l leg (0 skips)")
    ap.add_argument("--mixed", type=int, default=200_000,
                    help="mixed-class leg: quartets of the l<=3, degree 1..8 workload (0 skips it)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(batch, B, origin, target_s: float):
    from oracle import pyoracle

    kind = "reference" if pyoracle.available("reference") else "port"
    chk = pyoracle.load(kind)
    basis = pyoracle.make_basis(batch.shells.shell_list()[: 4 * 64], chk)
    threads = cpu_threads()
    f = chk.field6(B, origin)
    # calibrate on 2*threads quartets, then size the sample to ~target_s
    n0 = min(2 * threads, 64)
    t = time.perf_counter()
    chk.eri_batch(basis, batch.quartets[:n0], f, threads=threads)
    dt = time.perf_counter() - t
    per = dt / n0
    n = int(min(max(n0, target_s / max(per, 1e-9)), len(batch.quartets)))
    n = max(threads, (n // threads) * threads)
    sub = batch.shells.subset(np.arange(4 * n))
    basis = pyoracle.make_basis(sub.shell_list(), chk)
    q = np.arange(4 * n, dtype=np.uint32).reshape(n, 4)
    t = time.perf_counter()
    chk.eri_batch(basis, q, f, threads=threads)
    dt = time.perf_counter() - t
    return {
        "value": n / dt,
        "unit": "shell quartets/s",
        "cores": threads,
        "kind": kind,
        "sample": f"first {n} of the same (dd|dd) batch, OpenMP schedule(dynamic,1), {dt:.1f} s",
    }


def mixed_leg(n: int, tune: bool = True):
    """Mixed classes l <= 3 (weights (1 + sum l)^-2), contraction degree U{1..8}:
    the generator of the 2M-quartet workload at n quartets. One warmup call,
    one timed call, inputs and output resident on the device. With `tune`,
    the launch form of the generic kernel is autotuned on the same batch
    first (per-bucket grids, per-bucket work queues, one persistent launch)."""
    from giaoint_amd import _native, autotune, workloads

    batch = workloads.mixed_batch(n)
    sh = batch.shells
    coefs = _native.normalize_shells(sh.L, sh.nprim, sh.ncontr, sh.exps, sh.coefs)
    ctx = _native.Context(0)
    db = _native.DeviceBasis(ctx, sh.centers, sh.L, sh.nprim, sh.ncontr, sh.exps, coefs)
    total = int(db.quartet_sizes(batch.quartets).sum())
    out = ctx.alloc(total * 16)
    f = _native.field6(batch.B, batch.origin)
    flops = workloads.batch_flops(sh, batch.quartets)
    tuned = None
    if tune:
        ctx.free(out)  # the tuner holds its own output buffer
        tuned = autotune.tune_eri(db, batch.quartets, f, knobs=("eri_queue",), reps=1,
                                  workload=f"mixed l<=3 degree 1..8, {n} quartets")
        log(f"mixed: autotuned {tuned.values} " +
            " ".join(f"{t.value}={t.seconds:.3f}s" for t in tuned.trials))
        out = ctx.alloc(total * 16)
    db.eri_batch(batch.quartets, f, device_out=out)
    ctx.synchronize()
    t = time.perf_counter()
    db.eri_batch(batch.quartets, f, device_out=out)
    ctx.synchronize()
    dt = time.perf_counter() - t
    st = ctx.stats()
    ctx.free(out)
    log(f"mixed: {n} quartets in {dt:.2f} s")
    extra = {"autotune": tuned.as_dict()} if tuned else {}
    return {**extra,
        "workload": f"{n} quartets, classes l<=3 weighted (1+sum l)^-2, degree U{{1..8}}, seed 3, "
                    f"{st['primitive_quartets']} primitive quartets",
        "quartets_per_s": n / dt,
        "seconds": dt,
        "tflops_convention": flops / dt / 1e12,
        "launches": st["launches"],
    }


def high_l_leg(n_gggg: int, n_ffdd: int):
    """The high-angular-momentum precision workloads at throughput: (gg|gg)
    and (ff|dd), contraction degree 2, exponents 10^U(log10 0.05, log10 5e3).
    (ff|dd) runs on the class-specialized kernels through its (dd|ff) image,
    (gg|gg) on the generic kernel. A warmup call on the first quartets, then
    one timed call per class, output resident on the device."""
    from giaoint_amd import _native, workloads

    g, fd = workloads.precision_batches(n_gggg=n_gggg, n_ffdd=n_ffdd)
    out = {}
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
        db.eri_batch(batch.quartets[: min(n, 256)], f, device_out=dev)
        ctx.synchronize()
        t = time.perf_counter()
        db.eri_batch(batch.quartets, f, device_out=dev)
        ctx.synchronize()
        dt = time.perf_counter() - t
        ctx.free(dev)
        log(f"{name}: {n} quartets in {dt:.2f} s")
        out[name] = {"workload": batch.description, "quartets": n, "quartets_per_s": n / dt, "seconds": dt,
                     "tflops_convention": flops / dt / 1e12}
    return out


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


def cholesky_leg(rank_cap: int, block: int = 0, tune_rank: int = 0):
    """Streamed pivoted Cholesky of the 30-atom 4s3p2d1f set (1050 AOs), tau = 1e-8.

    Runs the device driver in steps of 1000 pivots (a progress line after
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
    # a device-resident consumer of the factors: J[mu,nu] = sum (mu nu|lam sig) D[lam,sig]
    rng = np.random.default_rng(0)
    dens = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    tj = time.perf_counter()
    ch.coulomb(dens)
    coulomb_s = time.perf_counter() - tj
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
        "pivot_blocks": blk["blocks"],
        "pivots_per_block": m / max(blk["blocks"], 1),
        "candidates_projected_per_block": blk["projected"] / max(blk["blocks"], 1),
        "projection": projection_roofline(proj_bytes, proj_flops, blk["block_ms"], unblocked_bytes),
    }
    ch.close()
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
    achieved = contract_flops / contract_launches / (avg_launch_ms * 1e-3) / 1e12
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
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel": "gh::fast::fast_transform_kernel<2,2,2,2> (E.R.E contraction)",
            "avg_launch_ms": avg_launch_ms,
            "flops_per_launch": contract_flops / contract_launches,
            "flop_convention": "8 per complex MAC of the cheaper dense Hermite-to-Cartesian contraction (per quartet "
                               f"{contract_flops / n:.4g}); the R0 part (12 per entry, {(flops - contract_flops) / n:.4g}"
                               " per quartet) runs in r0_kernel",
            "whole_path_tflops": flops / (ms_step * 1e-3) / 1e12,
            "executed_tflops": (executed - (flops - contract_flops)) / (contract_ms * 1e-3) / 1e12,
            "executed_note": "complex MACs the contraction executes (structural zeros of the Hermite products "
                             "skipped), 8 FLOP each",
            "fp64_issued_tflops": issued,
            "fp64_issued_frac": None if issued is None else issued / FP64_PEAK_TFLOPS,
            "fp64_issued_note": "FP64 operations the kernel issues (rocprofv3 SQ_INSTS_VALU_*_F64 per quartet) per "
                                "average launch: the FP64 pipe's utilisation. The convention counts the dense "
                                "contraction, whose structural zeros the kernel skips, so `frac` can exceed 1",
            "waves_per_workgroup": waves_used,
            "traffic_note": None if traffic is None else TRANSFORM_TRAFFIC_NOTE.format(
                b=TRANSFORM_HBM_BYTES_PER_QUARTET[waves_used], w=waves_used),
        },
        "checksum": float(np.abs(probe).sum()),
    }
    if tuned is not None:
        result["autotune"] = tuned.as_dict()
    result["fast_waves"] = fast_waves
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(batch, batch.B, batch.origin, args.cpu_seconds)
            result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
        except Exception as exc:  # the baseline is reported, never required
            result["cpu_baseline"] = {"error": str(exc)}
    ctx.free(out)
    del dbasis, ctx
    if args.mixed > 0:
        result["mixed_class"] = mixed_leg(args.mixed, tune=not args.no_autotune)
    if args.high_l_gggg > 0 or args.high_l_ffdd > 0:
        result["high_l"] = high_l_leg(args.high_l_gggg, args.high_l_ffdd)
    if args.cholesky_rank >= 0:
        result["cholesky"] = cholesky_leg(args.cholesky_rank, args.cholesky_block,
                                          0 if args.no_autotune else args.tune_cholesky_rank)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
