"""Launch-shape autotuner for g4acc.

For a workload and a batch size it measures, with the library's own HIP-event stage timing on
its stream, every candidate of each tunable of include/g4acc.h (g4acc_set_tuning):
  mtog_tw        frequencies per thread of the fused single-band M -> G kernel (M -> G stage)
  mtog_mfma      right half of that kernel on the matrix cores (complex64, Nc <= 16)
  g4_qg          momenta per G4 workgroup                                        (G4 stage)
  g4_tile_order  natural, XCD-aware super-tile or diagonal-strip tile schedule   (G4 stage)
  g4_mfma        rank-one G4 terms on the matrix cores or the vector ALUs         (G4 stage)
  g4_rank1       rank-one terms of the whole batch as one matrix-core GEMM          (G4 stage)
The stages are independent, so each is tuned on its own (the G4 knobs jointly). Results go to a
JSON table (dca_amd/tuning.json by default) keyed by the geometry, the precision and the batch
size; TpAccumulator applies the matching entry when it is constructed.

usage (on a GPU):  python -m dca_amd.autotune --config 2 [--walkers 64] [--precision fp32]
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

CANDIDATES = {
    "mtog_tw": [0, 1, 2, 4, 8],
    "mtog_mfma": [0, 1],
    "g4_qg": [0, 1, 2, 4],
    "g4_tile_order": [0, 1, 2],
    "g4_mfma": [0, 1],
    "g4_rank1": [0, 1],
}


def key(geom, precision: str, walkers: int) -> str:
    return (f"nb{geom.nb}_nc{geom.nc}_nw{geom.n_freq}_wex{'.'.join(map(str, geom.w_ex))}"
            f"_nq{len(geom.k_ex)}_ch{'.'.join(map(str, geom.channels))}_{precision}_W{walkers}")


def load_table(path: str = TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def lookup(geom, precision: str, walkers: int, path: str = TABLE) -> dict | None:
    entry = load_table(path).get(key(geom, precision, walkers))
    return None if entry is None else entry.get("params")


def _stage_times(acc, batches, reps: int):
    from dca_amd import g4acc
    acc.set_timing(True)
    lib, h = acc.lib, acc._h
    for k in range(reps):
        arr, n = batches[k % len(batches)]
        acc._check(lib.g4acc_accumulate_batch(h, arr, n, g4acc.DEVICE))
    mtog, g4, nb = acc.get_timing()
    acc.set_timing(False)
    return mtog / nb, g4 / nb


def tune(cfg: int, walkers: int, precision: str, reps: int = 4, log=print) -> dict:
    import torch
    from dca_amd import g4acc, synth
    from dca_amd.devpool import DevicePool

    wl = synth.workload(cfg)
    g = wl.geometry
    pool = wl.draw(walkers * 2)
    dp = DevicePool(torch, pool)
    batches = dp.batches(walkers)
    acc = g4acc.TpAccumulator(g, precision=precision, sum_mode="fp64", max_walkers=walkers,
                              max_vertices=wl.n_max, tuned=False)
    acc.resetAccumulation()
    _stage_times(acc, batches, 1)  # warm up
    measured = {"mtog": {}, "g4": {}}
    best_tw, best_mm, best_t = 0, 0, float("inf")
    # The matrix-core kernel has one shape (mtog_tw does not apply to it); where it does not
    # apply (complex128, Nc > 16) the setting falls back to the vector kernel and measures alike.
    for mm, tw in [(0, tw) for tw in CANDIDATES["mtog_tw"]] + [(1, 0)]:
        acc.set_tuning(mtog_tw=tw, mtog_mfma=mm)
        t, _ = _stage_times(acc, batches, reps)
        name = "mtog_mfma=1" if mm else f"mtog_tw={tw}"
        measured["mtog"][name] = round(t, 4)
        log(f"  M->G  {name}: {t:.3f} ms")
        if t < best_t:
            best_tw, best_mm, best_t = tw, mm, t
    acc.set_tuning(mtog_tw=best_tw, mtog_mfma=best_mm)
    best_g4, best_g4_t = (0, 0, 1, 0), float("inf")
    # The two rank-one forms exclude each other: g4_mfma forms them per walker on the matrix
    # cores inside the tile, g4_rank1 once per batch before it.
    for qg, order, mf, r1 in itertools.product(CANDIDATES["g4_qg"], CANDIDATES["g4_tile_order"],
                                               CANDIDATES["g4_mfma"], CANDIDATES["g4_rank1"]):
        if mf and r1:
            continue
        acc.set_tuning(g4_qg=qg, g4_tile_order=order, g4_mfma=mf, g4_rank1=r1)
        _, t = _stage_times(acc, batches, reps)
        name = f"g4_qg={qg},g4_tile_order={order},g4_mfma={mf},g4_rank1={r1}"
        measured["g4"][name] = round(t, 4)
        log(f"  G4    {name.replace(',', ' ')}: {t:.3f} ms")
        if t < best_g4_t:
            best_g4, best_g4_t = (qg, order, mf, r1), t
    acc.close()
    return {"params": {"mtog_tw": best_tw, "mtog_mfma": best_mm, "g4_qg": best_g4[0], "g4_tile_order": best_g4[1],
                       "g4_mfma": best_g4[2], "g4_rank1": best_g4[3]},
            "stage_ms": {"mtog": round(best_t, 4), "g4": round(best_g4_t, 4)},
            "measured": measured, "workload": cfg, "when": time.strftime("%Y-%m-%d %H:%M:%S")}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--config", type=int, nargs="+", default=[2])
    p.add_argument("--walkers", type=int, default=0)
    p.add_argument("--precision", nargs="+", default=["fp32", "fp64"])
    p.add_argument("--out", default=TABLE)
    args = p.parse_args()
    from dca_amd import synth
    table = load_table(args.out)
    for cfg in args.config:
        wl = synth.workload(cfg)
        for prec in args.precision:
            walkers = args.walkers or (wl.walkers if prec == "fp32" else 1)
            print(f"workload {cfg} {prec} walkers={walkers}", flush=True)
            entry = tune(cfg, walkers, prec, log=lambda m: print(m, flush=True))
            table[key(wl.geometry, prec, walkers)] = entry
            print(f"  best: {entry['params']} {entry['stage_ms']}", flush=True)
            with open(args.out, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
