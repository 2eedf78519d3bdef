#!/usr/bin/env python3
"""Benchmark: G4 measurements/s of the batched mixed-precision form vs the plain FP64 form of
the two-particle accumulator on one MI355X, with the FP64 CPU oracle timed on the host.

Workload (default): 1 band, 4x4 cluster, beta=10, 64 vertex frequencies, 4 k_ex, 3 w_ex, the
magnetic, charge and particle-particle channels, N per spin ~ Poisson(128) clipped to [64, 256],
M entries N(0, 1/N), sign +1 w.p. 0.9. A step is one batch of `walkers` measurements through the
whole path (M -> G -> G4) with the inputs already in HBM. The plain FP64 form is the same
algorithm with FP64 G and one measurement per launch sequence.

Prints one JSON line (rank 0).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G4 measurements/s: batched MFMA-FP32 vs plain FP64 at max rel err <=1e-6"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 (vector and matrix)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--config", type=int, default=2)
    p.add_argument("--walkers", type=int, default=0, help="walkers per batch (default: workload's)")
    p.add_argument("--pool", type=int, default=2, help="distinct batches cycled through")
    p.add_argument("--plain-steps", type=int, default=32)
    p.add_argument("--cpu-sample", type=int, default=8, help="measurements timed on the CPU oracle")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--copy-calibration", type=int, default=0,
                   help="bytes of a plain device-to-device copy run first (counter calibration)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(cfg, walkers):
    """HBM bytes per launch of the G4 stage from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of the same workload and batch size, or None."""
    path = os.path.join(ROOT, "profiles", f"g4_update_cfg{cfg}_hbm_traffic.txt")
    if not os.path.exists(path):
        return None, None
    traffic, w = None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            traffic = int(line.split(":")[1])
        if line.startswith("walkers_per_launch:"):
            w = int(line.split(":")[1])
    if traffic is None or w != walkers:
        return None, None
    return traffic, os.path.relpath(path, ROOT)


def run_batches(acc, batches, steps):
    from dca_amd import g4acc
    lib, h = acc.lib, acc._h
    for k in range(steps):
        arr, n = batches[k % len(batches)]
        rc = lib.g4acc_accumulate_batch(h, arr, n, g4acc.DEVICE)
        if rc != 0:
            acc._check(rc)


def cpu_oracle(wl, measurements, threads):
    import oracle
    oracle.build()
    accs = [oracle.OracleAccumulator(wl.geometry) for _ in range(threads)]
    shares = [measurements[i::threads] for i in range(threads)]

    def work(a, ms):
        for m in ms:
            a.accumulate(m)

    ts = [threading.Thread(target=work, args=(a, ms)) for a, ms in zip(accs, shares)]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    elapsed = time.perf_counter() - t0
    for a in accs[1:]:
        a.sumTo(accs[0])
    return elapsed, accs[0].get_G4()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(measurements shard across GPUs by walker: one process per GPU, "
                                   "G4 merged with one sum-reduce at the end)"}))
        sys.exit(2)
    import torch
    from dca_amd import g4acc, synth
    from dca_amd import geometry as geo
    from oracle import bruteforce as bf

    if args.copy_calibration:
        n = args.copy_calibration // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy of {n * 8} bytes done")
        del src, dst
    wl = synth.workload(args.config)
    g = wl.geometry
    W = args.walkers or wl.walkers
    log(f"workload {args.config}: {wl.description}; walkers/batch={W}")
    t0 = time.perf_counter()
    pool = wl.draw(W * args.pool)
    log(f"drew {len(pool)} measurements in {time.perf_counter() - t0:.1f}s")
    from dca_amd.devpool import DevicePool
    dp = DevicePool(torch, pool)
    batches = dp.batches(W)

    # ---- batched mixed precision form -------------------------------------------------------
    mixed = g4acc.TpAccumulator(g, precision="fp32", sum_mode="fp64", max_walkers=W,
                                max_vertices=wl.n_max)
    mixed.resetAccumulation()
    run_batches(mixed, batches, args.warmup)
    mixed.finalize()
    log("mixed warmup done")
    t0 = time.perf_counter()
    run_batches(mixed, batches, args.steps)
    mixed.finalize()
    t_mixed = time.perf_counter() - t0
    value = args.steps * W / t_mixed
    log(f"mixed: {args.steps} batches x {W} in {t_mixed:.3f}s -> {value:.1f} meas/s")

    # ---- plain FP64 form (one measurement per launch sequence) --------------------------------
    plain = g4acc.TpAccumulator(g, precision="fp64", sum_mode="fp64", max_walkers=1,
                                max_vertices=wl.n_max)
    plain.resetAccumulation()
    singles = dp.batches(1)
    run_batches(plain, singles, 2)
    plain.finalize()
    t0 = time.perf_counter()
    run_batches(plain, singles, args.plain_steps)
    plain.finalize()
    t_plain = time.perf_counter() - t0
    value_plain = args.plain_steps / t_plain
    log(f"plain: {args.plain_steps} measurements in {t_plain:.3f}s -> {value_plain:.1f} meas/s")

    # ---- stage timing with HIP events on the library's stream ---------------------------------
    mixed.set_timing(True)
    run_batches(mixed, batches, 4)
    mtog_ms, g4_ms, nb_ = mixed.get_timing()
    mixed.set_timing(False)
    mtog_ms /= nb_
    g4_ms /= nb_
    plain.set_timing(True)
    run_batches(plain, singles, 8)
    p_mtog, p_g4, p_n = plain.get_timing()
    plain.set_timing(False)
    p_mtog /= p_n
    p_g4 /= p_n
    log(f"stage ms per batch: mixed M->G {mtog_ms:.3f} G4 {g4_ms:.3f}; plain M->G {p_mtog:.3f} G4 {p_g4:.3f}")

    # Algorithmic HBM bytes of one G4-stage launch sequence (all channels): one FP64 complex
    # read + write of every G4 entry plus one read of every walker's G (complex64, both spins).
    n_ch = len(g.channels)
    g4_entries = g.g4_size()
    bytes_g4_stage = n_ch * 32 * g4_entries + W * 2 * g.D * g.D * 8
    achieved = bytes_g4_stage / (g4_ms * 1e-3) / 1e9
    # Algorithmic FP32 work of the same launch: per walker and G4 entry (k1, w1, k2, w2, q, v) the
    # complex multiply-adds of the single-band channels (tp_accumulator_cpu.hpp:491-661): rank-one
    # MAGNETIC and CHARGE, the crossed sum X0 Y0 + X1 Y1 shared by both, X0 Z1 + X1 Z0 for PP
    # (8 flops each). With batches the stage stops being HBM-bound; this is its compute rate.
    cfma = {geo.PARTICLE_HOLE_MAGNETIC: 1, geo.PARTICLE_HOLE_CHARGE: 1, geo.PARTICLE_PARTICLE_UP_DOWN: 2}
    n_cfma = sum(cfma.get(c, 0) for c in g.channels) + (
        2 if any(c in (geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE) for c in g.channels) else 0)
    flops_g4_stage = 8 * n_cfma * g4_entries * W if g.nb == 1 else None
    stage = "g4" if g4_ms >= mtog_ms else "mtog"
    r1 = bool((mixed.tuning or {}).get("g4_rank1"))
    g4_kernels = ("G4 update (uv_kernel" + (" + g4_rank1_kernel" if r1 else "") +
                  " + g4_tiled_kernel, one launch each per batch)")
    traffic, traffic_src = measured_traffic(args.config, W)

    # ---- accuracy against the FP64 CPU oracle, and the CPU baseline ---------------------------
    result_err = {}
    cpu = None
    if not args.no_cpu:
        S = max(1, min(args.cpu_sample, len(pool)))
        sample = pool[:S]
        threads = args.cpu_threads or min(16, os.cpu_count() or 1, S)
        log(f"CPU oracle on {S} measurements with {threads} threads ...")
        t_cpu, g4_ref = cpu_oracle(wl, sample, threads)
        cpu = {"value": S / t_cpu, "unit": "measurements/s", "cores": threads, "kind": "port",
               "sample": f"first {S} measurements of the workload, {threads} accumulators in "
                         f"{threads} threads (one per core), FP64, {t_cpu:.1f} s wall"}
        log(f"CPU oracle: {t_cpu:.2f}s -> {S / t_cpu:.3f} meas/s")
        for name, prec, mw in (("mixed", "fp32", W), ("plain_fp64", "fp64", 1)):
            acc = g4acc.TpAccumulator(g, precision=prec, sum_mode="fp64", max_walkers=mw,
                                      max_vertices=wl.n_max)
            acc.resetAccumulation()
            arr = (g4acc.Walker_c * S)(*dp.walkers[:S])
            acc._check(acc.lib.g4acc_accumulate_batch(acc._h, arr, S, g4acc.DEVICE))
            got = acc.get_G4()
            result_err[name] = max(bf.difference(r, x)["l_inf"] for r, x in zip(g4_ref, got))
            acc.close()
        log(f"G4 l_inf vs oracle: {result_err}")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "measurements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mixed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"cfg{args.config}: {wl.description}",
            "walkers_per_step": W,
            "g_precision": "complex64 (FP32 products, FP32 accumulation)",
            "g4_sum": "fp64",
            "plain_form": "complex128 G, FP64 G4, 1 measurement per launch sequence",
        },
        "plain_fp64_value": round(value_plain, 3),
        "speedup_vs_plain_fp64": round(value / value_plain, 3),
        "g4_linf_vs_oracle": result_err or None,
        "stage_ms_per_step": {"mixed_mtog": round(mtog_ms, 4), "mixed_g4": round(g4_ms, 4),
                              "plain_mtog": round(p_mtog, 4), "plain_g4": round(p_g4, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": g4_kernels,
                     "dominant_stage": stage,
                     "algorithmic_bytes_per_launch": bytes_g4_stage,
                     "traffic_source": traffic_src,
                     "compute": None if flops_g4_stage is None else {
                         "algorithmic_flops_per_launch": flops_g4_stage,
                         "achieved_tflops": round(flops_g4_stage / (g4_ms * 1e-3) / 1e12, 2),
                         "fp32_peak_tflops": FP32_PEAK_TFLOPS,
                         "frac": round(flops_g4_stage / (g4_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4)}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
