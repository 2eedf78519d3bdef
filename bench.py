#!/usr/bin/env python3
"""Benchmark: G4 measurements/s of the batched mixed-precision form vs the plain FP64 form of
the two-particle accumulator on one MI355X, with the FP64 CPU oracle timed on the host cores.

Headline workload (default, the largest single-GPU one): 1 band, 8x8 cluster, beta=40, 256
vertex frequencies, 1 k_ex, 1 w_ex, the charge channel, N per spin ~ U[800, 2000], M entries
N(0, 1/N), sign +1 w.p. 0.9, 16 walkers per batch. A step is one batch of `walkers` measurements
through the whole path (M -> G -> G4) with the inputs already in HBM. The plain FP64 form is the
same algorithm with FP64 G (complex128) and one measurement per launch sequence.

The other workloads (4x4, 6x6 and the three-band 8-site cluster) are timed after the headline
with fewer steps and reported under "workloads" (mixed and plain measurements/s, speedup, stage
times, mixed-vs-plain G4 error over the timed measurements).

The CPU baseline is the FP64 oracle (the reference CPU accumulator restated, its GEMMs through
the image's BLAS zgemm_ when that loads) on a bounded sample of the headline workload: one
accumulator per thread, as the reference's threaded solver runs them. The same sample is the
accuracy reference of the headline (l_inf of both GPU forms against the oracle).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import platform
import subprocess
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G4 measurements/s: batched MFMA-FP32 vs plain FP64 at max rel err <=1e-6"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 (vector with packed math, and matrix)
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 (vector and matrix)
HEADLINE = 5


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--config", type=int, default=HEADLINE, help="headline workload")
    p.add_argument("--extra", default="2,3,4", help="other workloads timed after the headline ('' for none)")
    p.add_argument("--extra-steps", type=int, default=6)
    p.add_argument("--walkers", type=int, default=0, help="walkers per batch (default: workload's)")
    p.add_argument("--pool", type=int, default=2, help="distinct batches cycled through")
    p.add_argument("--plain-steps", type=int, default=0, help="plain measurements timed (default: walkers)")
    p.add_argument("--cpu-threads", type=int, default=0,
                   help="oracle accumulators (threads) of the CPU baseline (default: by cores and RAM)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--copy-calibration", type=int, default=0,
                   help="bytes of a plain device-to-device copy run first (counter calibration)")
    p.add_argument("--sum-mode", default="fp64", choices=("fp64", "fp32_kahan"),
                   help="G4 running sum of the mixed form (the plain form always sums in FP64)")
    p.add_argument("--fixed-n", type=int, default=0,
                   help="every measurement of the headline gets this vertex count per spin instead "
                        "of the workload's law (stage times against N; at most the workload's N_max)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info():
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    mem_gb = None
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemAvailable:"):
                mem_gb = int(line.split()[1]) / 2 ** 20
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return {"nproc": cores, "cpu_model": model, "mem_available_gb": mem_gb}


def measured_traffic(cfg, walkers, key="traffic_bytes_per_launch"):
    """HBM bytes of one batch's G4 stage (key traffic_bytes_per_launch) or M -> G stage (key
    mtog_traffic_bytes_per_batch) from the committed rocprofv3 --pmc summary (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated) of the same workload and batch size, or None."""
    for name in (f"mtog_g4_cfg{cfg}_hbm_traffic.txt", f"g4_update_cfg{cfg}_hbm_traffic.txt"):
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            continue
        traffic, w = None, None
        for line in open(path):
            if line.startswith(key + ":"):
                traffic = int(line.split(":")[1])
            if line.startswith("walkers_per_launch:"):
                w = int(line.split(":")[1])
        if traffic is not None and w == walkers:
            return traffic, os.path.relpath(path, ROOT)
    return None, None


def run_batches(acc, batches, steps):
    from dca_amd import g4acc
    lib, h = acc.lib, acc._h
    for k in range(steps):
        arr, n = batches[k % len(batches)]
        rc = lib.g4acc_accumulate_batch(h, arr, n, g4acc.DEVICE)
        if rc != 0:
            acc._check(rc)


def mtog_flops(g, n):
    """FP operations of the M -> G algorithm this library runs, per slot of n vertices (complex
    multiply-add = 8, real x complex = 4; a site transform of Nc points as an L x L FFT,
    5 Nc log2 Nc, or as a dense Nc x Nc product, 8 Nc^2)."""
    import math
    nwE, nc, D = g.n_ext, g.nc, g.D
    L = int(round(math.sqrt(nc)))
    fft = L * L == nc and L in (2, 4, 8) and g.nb == 1
    site = 5 * nc * math.log2(nc) if fft else 8 * nc * nc
    return (4 * nwE * n * n            # S1: left NDFT (real M)
            + nwE * n * g.nb * site    # S2: left site transform
            + 8 * D * n * nwE          # S3: right NDFT
            + D * nwE * g.nb * site    # S4: right site transform
            + 12 * D * D)              # G = -G0 M~ G0 + beta G0 delta


def time_config(cfg, steps, warmup, walkers, pool_n, plain_steps, torch, keep_sample=0, fixed_n=0,
                sum_mode="fp64"):
    """Times the mixed (batched) and plain forms of one workload; returns a dict of results and,
    if keep_sample > 0, the first keep_sample measurements with both forms' G4 over them."""
    from dca_amd import g4acc, synth
    from dca_amd.devpool import DevicePool
    from oracle import bruteforce as bf
    wl = synth.workload(cfg)
    if fixed_n:
        if not 0 < fixed_n <= wl.n_max:
            raise SystemExit(f"--fixed-n must be in [1, {wl.n_max}] for workload {cfg}")
        wl = dataclasses.replace(wl, draw_n=lambda rs: fixed_n,
                                 description=f"{wl.description}; N fixed at {fixed_n}")
    g = wl.geometry
    W = walkers or wl.walkers
    log(f"workload {cfg}: {wl.description}; walkers/batch={W}")
    t0 = time.perf_counter()
    pool = wl.draw(W * pool_n)
    log(f"drew {len(pool)} measurements in {time.perf_counter() - t0:.1f}s")
    dp = DevicePool(torch, pool)
    batches = dp.batches(W)

    mixed = g4acc.TpAccumulator(g, precision="fp32", sum_mode=sum_mode, max_walkers=W,
                                max_vertices=wl.n_max)
    mixed.resetAccumulation()
    run_batches(mixed, batches, warmup)
    mixed.finalize()
    t0 = time.perf_counter()
    run_batches(mixed, batches, steps)
    mixed.finalize()
    t_mixed = time.perf_counter() - t0
    value = steps * W / t_mixed
    log(f"mixed: {steps} batches x {W} in {t_mixed:.3f}s -> {value:.2f} meas/s")

    plain = g4acc.TpAccumulator(g, precision="fp64", sum_mode="fp64", max_walkers=1,
                                max_vertices=wl.n_max)
    plain.resetAccumulation()
    singles = dp.batches(1)
    run_batches(plain, singles, 2)
    plain.finalize()
    n_plain = plain_steps or W
    t0 = time.perf_counter()
    run_batches(plain, singles, n_plain)
    plain.finalize()
    t_plain = time.perf_counter() - t0
    value_plain = n_plain / t_plain
    log(f"plain: {n_plain} measurements in {t_plain:.3f}s -> {value_plain:.2f} meas/s")

    # Stage times with HIP events on the library's stream.
    mixed.set_timing(True)
    run_batches(mixed, batches, min(4, max(2, steps)))
    mtog_ms, g4_ms, nb_ = mixed.get_timing()
    mixed.set_timing(False)
    plain.set_timing(True)
    run_batches(plain, singles, min(8, n_plain))
    p_mtog, p_g4, p_n = plain.get_timing()
    plain.set_timing(False)
    mtog_ms, g4_ms, p_mtog, p_g4 = mtog_ms / nb_, g4_ms / nb_, p_mtog / p_n, p_g4 / p_n
    log(f"stage ms: mixed M->G {mtog_ms:.3f} G4 {g4_ms:.3f} per batch; plain M->G {p_mtog:.3f} "
        f"G4 {p_g4:.3f} per measurement")

    # Mixed vs plain over the first batch of measurements (the plain form is pinned to the oracle
    # at ~1e-15 by the parity tests).
    acc_m = g4acc.TpAccumulator(g, precision="fp32", sum_mode=sum_mode, max_walkers=W,
                                max_vertices=wl.n_max)
    acc_p = g4acc.TpAccumulator(g, precision="fp64", sum_mode="fp64", max_walkers=1,
                                max_vertices=wl.n_max)
    res_g4 = {}
    for name, acc in (("mixed", acc_m), ("plain_fp64", acc_p)):
        acc.resetAccumulation()
        arr = (g4acc.Walker_c * W)(*dp.walkers[:W])
        acc._check(acc.lib.g4acc_accumulate_batch(acc._h, arr, W, g4acc.DEVICE))
        res_g4[name] = acc.get_G4()
    linf_mp = max(bf.difference(p, m)["l_inf"] for p, m in zip(res_g4["plain_fp64"], res_g4["mixed"]))
    sample = None
    if keep_sample:
        S = min(keep_sample, W)
        for name, acc in (("mixed", acc_m), ("plain_fp64", acc_p)):
            acc.resetAccumulation()
            arr = (g4acc.Walker_c * S)(*dp.walkers[:S])
            acc._check(acc.lib.g4acc_accumulate_batch(acc._h, arr, S, g4acc.DEVICE))
            res_g4[name] = acc.get_G4()
        sample = (pool[:S], res_g4)
    for a in (acc_m, acc_p, mixed, plain):
        a.close()
    del dp, batches, singles
    torch.cuda.empty_cache()

    n_mean = float(np.mean([n for m in pool for n in m.n]))
    out = {
        "workload": f"cfg{cfg}: {wl.description}",
        "walkers_per_step": W,
        "value": round(value, 3),
        "ms_per_step": round(1e3 * t_mixed / steps, 4),
        "plain_fp64_value": round(value_plain, 3),
        "speedup_vs_plain_fp64": round(value / value_plain, 3),
        "g4_linf_mixed_vs_plain_fp64": linf_mp,
        "stage_ms": {"mixed_mtog_per_batch": round(mtog_ms, 4), "mixed_g4_per_batch": round(g4_ms, 4),
                     "plain_mtog_per_meas": round(p_mtog, 4), "plain_g4_per_meas": round(p_g4, 4)},
    }
    ctx = {"wl": wl, "g": g, "W": W, "mtog_ms": mtog_ms, "g4_ms": g4_ms, "p_mtog": p_mtog, "p_g4": p_g4,
           "n_mean": n_mean, "pool": pool, "t_mixed": t_mixed}
    return out, ctx, sample


def roofline(cfg, ctx):
    """Roofline of the dominant stage of a mixed batch (by time)."""
    from dca_amd import geometry as geo
    g, W = ctx["g"], ctx["W"]
    mtog_ms, g4_ms = ctx["mtog_ms"], ctx["g4_ms"]
    D = g.D
    # G4 stage: one FP64 complex read + write of every G4 entry plus one read of every walker's
    # G (complex64, both spins).
    g4_bytes = len(g.channels) * 32 * g.g4_size() + W * 2 * D * D * 8
    g4_gbs = g4_bytes / (g4_ms * 1e-3) / 1e9
    # M -> G stage: compulsory bytes (M in, G out as complex64) and this library's flops.
    pool = ctx["pool"][:W]
    mtog_bytes = sum(8 * n * n + 8 * n for m in pool for n in m.n) + W * 2 * D * D * 8
    mtog_gbs = mtog_bytes / (mtog_ms * 1e-3) / 1e9
    flops = sum(mtog_flops(g, n) for m in pool for n in m.n)
    mtog_tf = flops / (mtog_ms * 1e-3) / 1e12
    g4_traffic, g4_src = measured_traffic(cfg, W)
    mtog_traffic, mtog_src = measured_traffic(cfg, W, "mtog_traffic_bytes_per_batch")
    hbm_frac, fp_frac = mtog_gbs / HBM_PEAK_GBS, mtog_tf / FP32_PEAK_TFLOPS
    if g4_ms >= mtog_ms:
        return {"bound": "hbm", "achieved": round(g4_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(g4_gbs / HBM_PEAK_GBS, 4), "traffic": g4_traffic,
                "kernel": "G4 update stage (uv_kernel + g4_tiled_kernel per batch)",
                "dominant_stage": "g4", "algorithmic_bytes_per_launch": g4_bytes,
                "traffic_source": g4_src,
                "other_stage": {"stage": "mtog", "ms": round(mtog_ms, 4), "achieved_tflops": round(mtog_tf, 2),
                                "frac_fp32": round(fp_frac, 4), "achieved_gbs": round(mtog_gbs, 1),
                                "frac_hbm": round(hbm_frac, 4), "traffic": mtog_traffic}}
    common = {"kernel": "M -> G stage (sort, phase, left and right M -> G kernels per batch)",
              "dominant_stage": "mtog", "traffic": mtog_traffic, "traffic_source": mtog_src,
              "algorithmic_flops_per_launch": flops, "algorithmic_bytes_per_launch": mtog_bytes,
              "other_stage": {"stage": "g4", "ms": round(g4_ms, 4), "achieved_gbs": round(g4_gbs, 1),
                              "frac_hbm": round(g4_gbs / HBM_PEAK_GBS, 4), "traffic": g4_traffic,
                              "algorithmic_bytes_per_launch": g4_bytes}}
    if fp_frac >= hbm_frac:
        return {"bound": "mfma", "achieved": round(mtog_tf, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(fp_frac, 4), "hbm_frac": round(hbm_frac, 4), **common}
    return {"bound": "hbm", "achieved": round(mtog_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(hbm_frac, 4), "compute_frac_fp32": round(fp_frac, 4), **common}


def cpu_baseline(wl, sample_ms, threads):
    """The oracle on the host cores: `threads` accumulators, one per thread (the reference's
    threading model), each taking its share of the sample. Returns (elapsed s, per-measurement
    single-thread seconds, G4 summed over the sample, GEMM name)."""
    import oracle
    oracle.build()
    gemm = oracle.use_blas(True)
    accs = [oracle.OracleAccumulator(wl.geometry) for _ in range(threads)]
    for a in accs:
        a.keep_intermediate(False)
    shares = [sample_ms[i::threads] for i in range(threads)]
    per = [[] for _ in range(threads)]

    def work(i):
        for m in shares[i]:
            t = time.perf_counter()
            accs[i].accumulate(m)
            per[i].append(time.perf_counter() - t)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    elapsed = time.perf_counter() - t0
    for a in accs[1:]:
        a.sumTo(accs[0])
    g4 = accs[0].get_G4()
    del accs
    oracle.use_blas(False)
    return elapsed, float(np.mean([x for p in per for x in p])), g4, gemm


def oracle_bytes_per_accumulator(g):
    """Host memory of one oracle accumulator: G and the NDFT intermediate (complex128, both spins)
    plus G4."""
    per_g = g.nb ** 2 * 2 * g.nc ** 2 * g.n_ext ** 2 * 16
    return 2 * per_g + len(g.channels) * g.g4_size() * 16


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(measurements shard across GPUs by walker: one process per GPU, "
                                   "G4 merged with one sum-reduce at the end)"}))
        sys.exit(2)
    import torch
    from oracle import bruteforce as bf

    if args.copy_calibration:
        n = args.copy_calibration // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy of {n * 8} bytes done")
        del src, dst

    host = host_info()
    # CPU sample of the headline: one measurement per accumulator thread, as many threads as the
    # host's CPU share and RAM allow (at most 16).
    from dca_amd import synth
    wl_h = synth.workload(args.config)
    per_acc = oracle_bytes_per_accumulator(wl_h.geometry)
    # The oracle accumulators take at most 200 GiB of host memory, 70 % of what is available.
    ram = min(host["mem_available_gb"] or 32, 200) * 2 ** 30
    threads = args.cpu_threads or max(1, min(16, host["nproc"], int(0.7 * ram // per_acc)))
    keep = 0 if args.no_cpu else threads

    head, ctx, sample = time_config(args.config, args.steps, args.warmup, args.walkers, args.pool,
                                    args.plain_steps, torch, keep_sample=keep, fixed_n=args.fixed_n,
                                    sum_mode=args.sum_mode)
    roof = roofline(args.config, ctx)

    extras = {}
    for c in [int(x) for x in args.extra.split(",") if x.strip()]:
        if c == args.config:
            continue
        res, cctx, _ = time_config(c, args.extra_steps, 2, 0, 2, 0, torch, sum_mode=args.sum_mode)
        res["roofline"] = roofline(c, cctx)
        extras[f"cfg{c}"] = res

    cpu, linf = None, None
    if sample is not None:
        ms, gpu_g4 = sample
        log(f"CPU oracle: {len(ms)} measurements of workload {args.config} on {threads} threads "
            f"({per_acc / 2 ** 30:.1f} GiB per accumulator) ...")
        elapsed, single, g4_ref, gemm = cpu_baseline(wl_h, ms, threads)
        log(f"CPU oracle: {elapsed:.1f}s wall, {single:.2f}s per measurement per thread ({gemm})")
        cpu = {"value": round(len(ms) / elapsed, 5), "unit": "measurements/s", "cores": threads,
               "kind": "port",
               "sample": f"first {len(ms)} measurements of the headline workload, one FP64 oracle "
                         f"accumulator per thread on {threads} threads, {elapsed:.1f} s wall",
               "single_thread_s_per_measurement": round(single, 3),
               "gemm": gemm, "nproc": host["nproc"], "cpu_model": host["cpu_model"]}
        linf = {name: max(bf.difference(r, x)["l_inf"] for r, x in zip(g4_ref, gpu_g4[name]))
                for name in ("mixed", "plain_fp64")}
        log(f"G4 l_inf vs oracle: {linf}")

    W = ctx["W"]
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "measurements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": head["workload"],
            "walkers_per_step": W,
            "g_precision": "complex64 (FP32 products, FP32 accumulation)",
            "g4_sum": args.sum_mode,
            "plain_form": "complex128 G, FP64 G4, 1 measurement per launch sequence",
        },
        "plain_fp64_value": head["plain_fp64_value"],
        "speedup_vs_plain_fp64": head["speedup_vs_plain_fp64"],
        "g4_linf_vs_oracle": linf,
        "g4_linf_mixed_vs_plain_fp64": head["g4_linf_mixed_vs_plain_fp64"],
        "stage_ms": head["stage_ms"],
        "roofline": roof,
        "cpu_baseline": cpu,
        "workloads": extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
