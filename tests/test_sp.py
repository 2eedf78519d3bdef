"""Single-particle accumulator (SpAccumulator, sp_accumulator.hpp / sp_accumulator_gpu.hpp).

CPU: the restated delayed NFFT (oracle/sp_oracle.py) is pinned by what it approximates, the exact
transform (1/Nc) sum_ij M_ij exp(i w (tau_i - tau_j)) binned by (b_i, b_j, r_i - r_j): it agrees
to the NFFT's own approximation error, ~6e-9 max-normalised from 32 frequencies on (the cubic
interpolation of the window on a grid of 32 points per coarse step, which does not shrink with the
frequency count; ~3e-7 for 8 frequencies near the window's band edge), and planted mistakes (a
flipped time difference, the band pair swapped) are far from it.
GPU: the HIP accumulator (exact FP64 transform) against the exact transform at 1e-12 and against
the restated NFFT at 1e-7 (its approximation error, with margin).
"""
import math

import numpy as np
import pytest

from dca_amd import geometry as geo
from dca_amd import synth
from oracle import bruteforce as bf
from oracle import sp_oracle as so


def _wl(nb, L, n_range=(2, 12), seed=500):
    return synth.tiny_workload(nb=nb, L=L, n_freq=4, n_range=n_range, seed_cfg=seed)


@pytest.mark.parametrize("nb,L,n_sp,tol", [(1, 2, 8, 1e-6), (2, 2, 32, 2e-8), (1, 3, 64, 2e-8), (1, 2, 128, 2e-8)])
def test_nfft_restatement_approximates_the_exact_transform(nb, L, n_sp, tol):
    wl = _wl(nb, L)
    g = wl.geometry
    ms = wl.draw(3)
    sp = so.SpOracle(nb, g.nc, g.beta, n_sp, g.cluster.rsub)
    for m in ms:
        sp.accumulate(m, m.sign)
    got = sp.get_sign_times_M_r_w()
    exact = so.m_r_w_exact(nb, g.nc, g.beta, n_sp, g.cluster.rsub, ms)
    err = bf.difference(exact, got)["l_inf"]
    assert err < tol, err
    # spin sectors are not mixed
    assert np.all(got[:, 0, :, 1] == 0) and np.all(got[:, 1, :, 0] == 0)


def test_exact_transform_detects_planted_mistakes():
    wl = _wl(2, 3, seed=501)
    g = wl.geometry
    ms = wl.draw(2)
    sp = so.SpOracle(2, g.nc, g.beta, 64, g.cluster.rsub)
    for m in ms:
        sp.accumulate(m, m.sign)
    got = sp.get_sign_times_M_r_w()
    exact = so.m_r_w_exact(2, g.nc, g.beta, 64, g.cluster.rsub, ms)
    assert bf.difference(exact, got)["l_inf"] < 2e-8
    flipped = [synth.Measurement(n=m.n, M=[x.T.copy(order="F") for x in m.M], tau=m.tau, site=m.site,
                                 band=m.band, sign=m.sign) for m in ms]   # (tau_j - tau_i)
    wrong = so.m_r_w_exact(2, g.nc, g.beta, 64, g.cluster.rsub, flipped)
    assert bf.difference(wrong, got)["l_inf"] > 1e-3
    swapped = np.swapaxes(exact, 0, 2)                                     # (b2, b1)
    assert bf.difference(swapped, got)["l_inf"] > 1e-3


def test_squared_accumulation():
    wl = _wl(1, 2, seed=502)
    g = wl.geometry
    ms = wl.draw(2)
    sp = so.SpOracle(1, g.nc, g.beta, 64, g.cluster.rsub, accumulate_m_sqr=True)
    for m in ms:
        sp.accumulate(m, m.sign)
    exact = so.m_r_w_exact(1, g.nc, g.beta, 64, g.cluster.rsub, ms, squared=True)
    assert bf.difference(exact, sp.get_sign_times_M_r_w_sqr())["l_inf"] < 2e-8


def test_site_subtraction_tables():
    for cl in (geo.square_cluster(2), geo.square_cluster(3), geo.cluster_2d((2, 2), (2, -2))):
        nc = cl.nc
        for a in range(nc):
            assert cl.rsub[a, a] == np.argmin(np.abs(cl.r).sum(axis=1))   # r_a - r_a = origin
            for b in range(nc):
                d = cl.r[b] - cl.r[a]
                # exp(i k.(r_b - r_a)) is the same for the folded site
                assert np.allclose(np.exp(1j * cl.k @ d), cl.dft[:, cl.rsub[a, b]])


@pytest.mark.gpu
@pytest.mark.parametrize("nb,L,n_sp", [(1, 2, 64), (3, 2, 32), (1, 4, 64)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gpu_sp_accumulator(nb, L, n_sp, precision):
    """The HIP single-particle accumulator (exact FP64 transform, beside the G4 path of either
    precision) against the exact transform and the restated NFFT, in batches, with sumTo."""
    from dca_amd import g4acc
    wl = _wl(nb, L, n_range=(3, 20), seed=503)
    g = wl.geometry
    ms = wl.draw(7)
    a = g4acc.TpAccumulator(g, precision=precision, max_walkers=3, max_vertices=wl.n_max)
    a.enable_sp(n_sp, accumulate_m_sqr=True)
    a.resetAccumulation()
    a.accumulate_batch(ms[:4])
    b = g4acc.TpAccumulator(g, precision=precision, max_walkers=3, max_vertices=wl.n_max)
    b.enable_sp(n_sp, accumulate_m_sqr=True)
    b.resetAccumulation()
    b.accumulate_batch(ms[4:])
    b.sumTo(a)
    got = a.get_sign_times_M_r_w()
    exact = so.m_r_w_exact(nb, g.nc, g.beta, n_sp, g.cluster.rsub, ms)
    assert bf.difference(exact, got)["l_inf"] < 1e-12
    exact2 = so.m_r_w_exact(nb, g.nc, g.beta, n_sp, g.cluster.rsub, ms, squared=True)
    assert bf.difference(exact2, a.get_sign_times_M_r_w(squared=True))["l_inf"] < 1e-12
    ref = so.SpOracle(nb, g.nc, g.beta, n_sp, g.cluster.rsub)
    for m in ms:
        ref.accumulate(m, m.sign)
    err = bf.difference(ref.get_sign_times_M_r_w(), got)["l_inf"]
    print(f"sp nb={nb} L={L} n_sp={n_sp}: HIP vs restated NFFT l_inf={err:.3e}", flush=True)
    assert err < 1e-7
    a.resetAccumulation()
    assert np.all(a.get_sign_times_M_r_w() == 0)
