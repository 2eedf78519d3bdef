"""The five measured workloads at their full shapes (GPU).

Workloads 1 and 2 are compared with the FP64 oracle directly. For all five, properties that do
not need the oracle at full size:
  * M = 0 closed forms (oracle/bruteforce.py: g4_free) for every channel of the workload;
  * a batched launch equals the same walkers accumulated one by one;
  * the mixed-precision form agrees with the FP64 form to 1e-6 (max-normalised l_inf).
"""
import numpy as np
import pytest

from dca_amd import g4acc, synth
from oracle import bruteforce as bf

pytestmark = pytest.mark.gpu


def linf(ref, x):
    return bf.difference(ref, x)["l_inf"]


def acc_for(wl, precision, walkers, sum_mode="fp64"):
    a = g4acc.TpAccumulator(wl.geometry, precision=precision, sum_mode=sum_mode,
                            max_walkers=walkers, max_vertices=wl.n_max)
    a.resetAccumulation()
    return a


@pytest.mark.parametrize("cfg,count", [(1, 8), (2, 1)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_workload_vs_oracle(oracle_lib, cfg, count, precision, tol):
    wl = synth.workload(cfg)
    ms = wl.draw(count)
    ref = oracle_lib.OracleAccumulator(wl.geometry)
    for m in ms:
        ref.accumulate(m)
    a = acc_for(wl, precision, walkers=count)
    a.accumulate_batch(ms)
    for r, x in zip(ref.get_G4(), a.get_G4()):
        err = linf(r, x)
        print(f"workload {cfg} {precision} l_inf={err:.3e}", flush=True)
        assert err < tol


@pytest.mark.parametrize("cfg", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_workload_free_closed_form(cfg, precision, tol):
    wl = synth.workload(cfg)
    g = wl.geometry
    m = wl.draw(1)[0]
    m.M = [np.zeros_like(x) for x in m.M]
    a = acc_for(wl, precision, walkers=1)
    a.accumulate(m)
    for ch, x in zip(g.channels, a.get_G4()):
        ref = bf.g4_free(g, ch, m.sign).ravel(order="F")
        err = linf(ref, x)
        print(f"workload {cfg} {precision} channel {ch} M=0 l_inf={err:.3e}", flush=True)
        assert err < tol
        del ref


@pytest.mark.parametrize("cfg,count", [(2, 6), (3, 3), (4, 4), (5, 2)])
def test_workload_batched_and_mixed(cfg, count):
    wl = synth.workload(cfg)
    ms = wl.draw(count)
    seq = acc_for(wl, "fp64", walkers=1)
    for m in ms:
        seq.accumulate(m)
    seq_g4 = seq.get_G4()
    seq.close()
    bat = acc_for(wl, "fp64", walkers=count)
    bat.accumulate_batch(ms)
    for r, x in zip(seq_g4, bat.get_G4()):
        err = linf(r, x)
        print(f"workload {cfg} batched vs sequential l_inf={err:.3e}", flush=True)
        assert err < 1e-13
    bat.close()
    mixed = acc_for(wl, "fp32", walkers=count)
    mixed.accumulate_batch(ms)
    for r, x in zip(seq_g4, mixed.get_G4()):
        err = linf(r, x)
        print(f"workload {cfg} mixed vs fp64 l_inf={err:.3e}", flush=True)
        assert err < 1e-6
