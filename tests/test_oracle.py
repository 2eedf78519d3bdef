"""Pins the FP64 CPU oracle (oracle/tp_oracle.cpp) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement (a dropped term, a wrong
sign, index or conjugation, a transposed band pair) fails it:
  * the defining NDFT double sum the reference's own tests use
    (test/unit/.../accumulation/single_sector_accumulation_test.hpp:226-254), at their tolerance
    of 100 eps (test/unit/.../tp/ndft/cached_ndft_gpu_test.cpp:68);
  * the rank-1 M of the reference's cached_ndft_test.hpp:100-118, whose NDFT factorises;
  * a single vertex (closed form) and the full direct M -> G definition;
  * M = 0 closed forms of every channel, for one and three bands, with non-zero transfers;
  * linearity in the sign, additivity, sumTo, the empty configuration;
  * the conjugation symmetry of a real M's NDFT.
"""
import math

import numpy as np
import pytest

from dca_amd import geometry as geo
from dca_amd import synth
from oracle import bruteforce as bf

EPS = np.finfo(np.float64).eps


def _wl(nb=1, L=2, n_freq=4, w_ex=(0, 1), k_ex=((0.0, 0.0), (math.pi, math.pi)), n_range=(1, 7), **kw):
    return synth.tiny_workload(nb=nb, L=L, n_freq=n_freq, w_ex=w_ex, k_ex=k_ex, n_range=n_range, **kw)


@pytest.mark.parametrize("nb", [1, 2, 3])
def test_ndft_matches_definition(oracle_lib, nb):
    wl = _wl(nb=nb)
    acc = oracle_lib.OracleAccumulator(wl.geometry)
    for meas in wl.draw(3):
        acc.compute_G(meas)
        mr = acc.get_Mrrww()
        for s in range(2):
            ref = bf.ndft_direct(wl.geometry, meas, s)
            assert bf.difference(ref, mr[:, :, :, :, s])["l_inf"] < 100 * EPS


def test_ndft_rank_one_factorises(oracle_lib):
    # cached_ndft_test.hpp:100-118: tau in [-0.5, 0.5), M_ij = sin(2 pi t_i) sin(6 pi t_j).
    wl = _wl(nb=2, L=2, n_freq=6, w_ex=(0,), k_ex=((0.0, 0.0),))
    g = wl.geometry
    rs = synth.Stream(12345)
    n = 31
    u = rs.uniform(3 * n)
    tau = u[:n] - 0.5
    band = np.minimum((u[n:2 * n] * g.nb).astype(np.int32), g.nb - 1)
    site = np.minimum((u[2 * n:] * g.nc).astype(np.int32), g.nc - 1)
    M = np.asfortranarray(np.outer(np.sin(2 * np.pi * tau), np.sin(6 * np.pi * tau)))
    meas = synth.Measurement(n=[n, 0], M=[M, np.zeros((0, 0))], tau=[tau, np.zeros(0)],
                             site=[site, np.zeros(0, np.int32)], band=[band, np.zeros(0, np.int32)],
                             sign=1.0)
    acc = oracle_lib.OracleAccumulator(g)
    acc.compute_G(meas)
    mr = acc.get_Mrrww()[:, :, :, :, 0]
    w = geo.vertex_frequencies(g.n_ext, g.beta)
    a = np.zeros((g.nc, g.nb, g.n_ext), complex)
    b = np.zeros((g.nc, g.nb, g.n_ext), complex)
    for i in range(n):
        a[site[i], band[i]] += np.sin(2 * np.pi * tau[i]) * np.exp(1j * tau[i] * w)
        b[site[i], band[i]] += np.sin(6 * np.pi * tau[i]) * np.exp(-1j * tau[i] * w)
    ref = np.einsum("rbw,sdv->rsbdwv", a, b)
    assert bf.difference(ref, mr)["l_inf"] < 100 * EPS


def test_single_vertex_closed_form(oracle_lib):
    wl = _wl(nb=1, L=3, n_freq=4, w_ex=(0, 2), k_ex=((0.0, 0.0),))
    g = wl.geometry
    tau, site, m = 0.7312, 2, -0.83
    meas = synth.Measurement(n=[1, 0], M=[np.full((1, 1), m, order="F"), np.zeros((0, 0))],
                             tau=[np.array([tau]), np.zeros(0)], site=[np.array([site], np.int32),
                                                                        np.zeros(0, np.int32)],
                             band=[np.array([0], np.int32), np.zeros(0, np.int32)], sign=1.0)
    acc = oracle_lib.OracleAccumulator(g)
    acc.compute_G(meas)
    G = acc.get_G()
    w = geo.vertex_frequencies(g.n_ext, g.beta)
    r = g.cluster.r[site]
    kr = g.cluster.k @ r
    mt = m * np.exp(1j * (kr[:, None, None, None] - kr[None, :, None, None]) +
                    1j * tau * (w[None, None, :, None] - w[None, None, None, :])) / g.nc
    g0 = g.g0[0, 0, 0]  # (k, w)
    ref = -g0[:, None, :, None] * mt * g0[None, :, None, :]
    for k in range(g.nc):
        for i in range(g.n_ext):
            ref[k, k, i, i] += g.beta * g0[k, i]
    assert bf.difference(ref, G[0, 0, 0])["l_inf"] < 1e-13
    # Spin 1 has no vertex: G = beta G0 delta exactly.
    free = np.zeros_like(ref)
    for k in range(g.nc):
        for i in range(g.n_ext):
            free[k, k, i, i] = g.beta * g0[k, i]
    assert np.array_equal(free, G[0, 0, 1])


@pytest.mark.parametrize("nb,L", [(1, 2), (1, 3), (3, 2)])
def test_G_matches_definition(oracle_lib, nb, L):
    wl = _wl(nb=nb, L=L, k_ex=((0.0, 0.0),))
    acc = oracle_lib.OracleAccumulator(wl.geometry)
    for meas in wl.draw(2):
        acc.compute_G(meas)
        assert bf.difference(bf.g_direct(wl.geometry, meas), acc.get_G())["l_inf"] < 1e-13


def test_real_M_conjugation_symmetry(oracle_lib):
    # For real M, M~(-w1, -w2) = conj M~(w1, w2) (used by cached_ndft_cpu_test.cpp:92-103).
    wl = _wl(nb=1, L=2, n_freq=6)
    acc = oracle_lib.OracleAccumulator(wl.geometry)
    acc.compute_G(wl.draw(1)[0])
    mr = acc.get_Mrrww()
    flipped = mr[..., ::-1, ::-1]
    assert bf.difference(mr, np.conj(flipped))["l_inf"] < 100 * EPS


@pytest.mark.parametrize("nb,L", [(1, 2), (1, 3), (3, 2)])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_free_closed_forms(oracle_lib, nb, L, sign):
    k_ex = ((0.0, 0.0), (math.pi, math.pi)) if L == 2 else ((0.0, 0.0), (2 * math.pi / 3, 0.0))
    wl = _wl(nb=nb, L=L, n_freq=4, w_ex=(0, 1, 2), k_ex=k_ex,
             channels=[geo.PARTICLE_HOLE_TRANSVERSE, geo.PARTICLE_HOLE_MAGNETIC,
                       geo.PARTICLE_HOLE_CHARGE, geo.PARTICLE_PARTICLE_UP_DOWN])
    g = wl.geometry
    meas = wl.draw(1)[0]
    meas.M = [np.zeros_like(x) for x in meas.M]
    meas.sign = sign
    acc = oracle_lib.OracleAccumulator(g)
    acc.accumulate(meas)
    g4 = acc.get_G4()
    for c, ch in enumerate(g.channels):
        ref = bf.g4_free(g, ch, sign).ravel(order="F")
        assert np.abs(ref).max() > 0
        assert bf.difference(ref, g4[c])["l_inf"] < 1e-14, geo.CHANNEL_NAMES[ch]


def test_closed_forms_detect_band_transposition(oracle_lib):
    """The three-band G0 is not band-symmetric, so the closed forms distinguish a transposed
    band pair: swapping (b1,b3) in the expected pp G4 must be far from the oracle."""
    wl = _wl(nb=3, L=2, n_freq=4, w_ex=(0, 1), channels=[geo.PARTICLE_PARTICLE_UP_DOWN])
    g = wl.geometry
    assert np.abs(g.g0 - np.swapaxes(g.g0, 0, 1)).max() > 1e-2
    meas = wl.draw(1)[0]
    meas.M = [np.zeros_like(x) for x in meas.M]
    acc = oracle_lib.OracleAccumulator(g)
    acc.accumulate(meas)
    ref = bf.g4_free(g, geo.PARTICLE_PARTICLE_UP_DOWN, meas.sign)
    wrong = np.swapaxes(ref, 0, 2).ravel(order="F")
    assert bf.difference(wrong, acc.get_G4()[0])["l_inf"] > 1e-3


@pytest.mark.parametrize("nb", [1, 3])
def test_linearity_and_sum_to(oracle_lib, nb):
    wl = _wl(nb=nb)
    g = wl.geometry
    m1, m2 = wl.draw(2)
    a = oracle_lib.OracleAccumulator(g)
    m1.sign = 1.0
    a.accumulate(m1)
    plus = a.get_G4()
    a.resetAccumulation()
    m1.sign = -1.0
    a.accumulate(m1)
    minus = a.get_G4()
    for p, m in zip(plus, minus):
        assert np.max(np.abs(p + m)) == 0.0
    a.resetAccumulation()
    a.accumulate(m1)
    a.accumulate(m1)
    for d, m in zip(a.get_G4(), minus):
        assert np.allclose(d, 2 * m, rtol=0, atol=1e-13 * np.abs(m).max())
    # sumTo == sequential accumulation; the source is cleared.
    b, c, seq = (oracle_lib.OracleAccumulator(g) for _ in range(3))
    b.accumulate(m1)
    c.accumulate(m2)
    seq.accumulate(m1)
    seq.accumulate(m2)
    b.sumTo(c)
    for x, y in zip(c.get_G4(), seq.get_G4()):
        assert np.allclose(x, y, rtol=0, atol=1e-13 * np.abs(y).max())
    with pytest.raises(RuntimeError):
        b.get_G4()


def test_empty_configuration_is_noop(oracle_lib):
    wl = _wl()
    a = oracle_lib.OracleAccumulator(wl.geometry)
    empty = synth.Measurement(n=[0, 0], M=[np.zeros((0, 0))] * 2, tau=[np.zeros(0)] * 2,
                              site=[np.zeros(0, np.int32)] * 2, band=[np.zeros(0, np.int32)] * 2,
                              sign=1.0)
    assert a.accumulate(empty) == 0.0
    assert all(np.all(x == 0) for x in a.get_G4())
