"""The five measured workloads at their full shapes (GPU).

Workload 1 is compared with the FP64 oracle over all of its 1000 measurements, workload 2 over
one. For all five, properties that do not need the oracle at full size:
  * M = 0 closed forms (oracle/bruteforce.py: g4_free) for every channel of the workload;
  * a batched launch equals the same walkers accumulated one by one;
  * the mixed-precision form agrees with the FP64 form to 1e-6 (max-normalised l_inf), on a few
    measurements and accumulated over each workload's full measurement count (10^4 for workload
    2), with the sum mode the workload names (FP32 + Kahan for workload 4). The FP64 GPU form is
    itself pinned to the oracle at ~1e-15 on every case of test_gpu_parity.py.
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


@pytest.mark.parametrize("cfg,count", [(1, 1000), (2, 1)])
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


# (workload, measurements, sum mode of the mixed form): each workload's full measurement count
ACCUMULATED = [(2, 10000, "fp64"), (3, 64 * 16, "fp64"), (4, 32 * 16, "fp32_kahan"), (5, 16 * 8, "fp64")]


@pytest.mark.parametrize("cfg,count,sum_mode", ACCUMULATED)
def test_accumulated_mixed_error(cfg, count, sum_mode):
    """The batched mixed-precision form against the plain FP64 form over a whole workload: the
    error that builds up in the running G4 sum over all measurements (sign-weighted, signs
    +1 w.p. 0.9) stays within the 1e-6 accuracy gate."""
    import time
    wl = synth.workload(cfg)
    W = wl.walkers
    plain = acc_for(wl, "fp64", walkers=1)
    mixed = acc_for(wl, "fp32", walkers=W, sum_mode=sum_mode)
    t0 = time.perf_counter()
    for chunk in wl.draw_chunks(count, W):
        mixed.accumulate_batch(chunk)
        plain.accumulate_batch(chunk)
        mixed.synchronizeCopy()
        plain.synchronizeCopy()
    errs = [linf(r, x) for r, x in zip(plain.get_G4(), mixed.get_G4())]
    print(f"workload {cfg}: {count} measurements, mixed ({sum_mode} G4 sum) vs plain FP64, "
          f"accumulated l_inf per channel = {['%.3e' % e for e in errs]} "
          f"({time.perf_counter() - t0:.1f} s)", flush=True)
    assert max(errs) < 1e-6
