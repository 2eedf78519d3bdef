"""Pins the FP64 CPU oracle (oracle/tp_oracle.cpp) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement (a dropped term, a wrong
sign, index or conjugation, a transposed band pair) fails it:
  * the defining NDFT double sum the reference's own tests use
    (test/unit/.../accumulation/single_sector_accumulation_test.hpp:226-254), at their tolerance
    of 100 eps (test/unit/.../tp/ndft/cached_ndft_gpu_test.cpp:68);
  * the rank-1 M of the reference's cached_ndft_test.hpp:100-118, whose NDFT factorises;
  * a single vertex (closed form) and the full direct M -> G definition;
  * M = 0 closed forms of every channel, for one and three bands, with non-zero transfers, with a
    spin-independent and a Zeeman-split G0 (the latter keeps the magnetic spin-difference term and
    the s / -s pairings of the transverse and pp channels from vanishing or coinciding);
  * one vertex per spin: G in closed form (diagonal plus a non-symmetric rank-one part) and G4 by
    the channel definitions, which sees the argument order of every off-diagonal G read;
  * planted mistakes in those closed forms (argument order, spin pairing, spin weights, the pp
    comment's transposed order) are far from the oracle, so the checks can fail;
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


ALL_CHANNELS = [geo.PARTICLE_HOLE_TRANSVERSE, geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE,
                geo.PARTICLE_PARTICLE_UP_DOWN]


@pytest.mark.parametrize("nb,L", [(1, 2), (1, 3), (3, 2)])
@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("zeeman", [0.0, 0.37])
def test_free_closed_forms(oracle_lib, nb, L, sign, zeeman):
    k_ex = ((0.0, 0.0), (math.pi, math.pi)) if L == 2 else ((0.0, 0.0), (2 * math.pi / 3, 0.0))
    wl = _wl(nb=nb, L=L, n_freq=4, w_ex=(0, 1, 2), k_ex=k_ex, channels=ALL_CHANNELS, zeeman=zeeman)
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


def _one_vertex_each(g, rs):
    """One vertex per spin at distinct (tau, site, band) with distinct M values, of a size (~Nc)
    that makes the rank-one part of G comparable with beta G0."""
    verts, n, Ms, taus, sites, bands = [], [], [], [], [], []
    for s in range(2):
        u = rs.uniform(4)
        tau, site, band, m = u[0] * g.beta, int(u[1] * g.nc), int(u[2] * g.nb), g.nc * (1.0 + u[3])
        verts.append((tau, site, band, m))
        n.append(1)
        Ms.append(np.full((1, 1), m, order="F"))
        taus.append(np.array([tau]))
        sites.append(np.array([site], np.int32))
        bands.append(np.array([band], np.int32))
    return verts, synth.Measurement(n=n, M=Ms, tau=taus, site=sites, band=bands, sign=-1.0)


SINGLE_VERTEX_CASES = {
    # name: (nb, L, w_ex, k_ex)
    "1band_2x2": (1, 2, (0, 1, 2), ((0.0, 0.0), (math.pi, 0.0))),
    "1band_3x3": (1, 3, (0, 2), ((2 * math.pi / 3, 0.0), (0.0, 2 * math.pi / 3))),
    "1band_3x3_negative_wex": (1, 3, (-1,), ((2 * math.pi / 3, 2 * math.pi / 3),)),
    "3band_2x2": (3, 2, (0, 1), ((0.0, 0.0), (math.pi, math.pi))),
}


def _single_vertex_case(oracle_lib, name, seed):
    nb, L, w_ex, k_ex = SINGLE_VERTEX_CASES[name]
    wl = _wl(nb=nb, L=L, n_freq=4, w_ex=w_ex, k_ex=k_ex, channels=ALL_CHANNELS, zeeman=0.41)
    g = wl.geometry
    verts, meas = _one_vertex_each(g, synth.Stream(seed))
    acc = oracle_lib.OracleAccumulator(g)
    acc.accumulate(meas)
    return g, verts, meas, acc


@pytest.mark.parametrize("name", list(SINGLE_VERTEX_CASES))
def test_single_vertex_G_closed_form(oracle_lib, name):
    g, verts, meas, acc = _single_vertex_case(oracle_lib, name, 7)
    ref = bf.g_single_vertex(g, verts)
    assert bf.difference(ref, acc.get_G())["l_inf"] < 1e-13
    # the rank-one part is not symmetric, so a transposed read of it would differ
    r1 = (ref - bf.g_single_vertex(g, [None, None]))[:, :, 0]  # (b1, b2, k1, k2, w1, w2)
    assert np.abs(r1).max() > 0
    r1t = np.transpose(r1, (1, 0, 3, 2, 5, 4))
    assert np.abs(r1 - r1t).max() > 0.1 * np.abs(r1).max()


@pytest.mark.parametrize("name", list(SINGLE_VERTEX_CASES))
@pytest.mark.parametrize("seed", [11, 12])
def test_single_vertex_g4_definition(oracle_lib, name, seed):
    """The oracle's accumulate of one vertex per spin (Zeeman-split G0, non-zero transfers)
    against the channel definitions evaluated on the closed-form G."""
    g, verts, meas, acc = _single_vertex_case(oracle_lib, name, seed)
    G = bf.g_single_vertex(g, verts)
    for c, ch in enumerate(g.channels):
        ref = bf.g4_definition(g, G, ch, meas.sign).ravel(order="F")
        err = bf.difference(ref, acc.get_G4()[c])["l_inf"]
        assert err < 1e-13, (geo.CHANNEL_NAMES[ch], err)


def test_free_closed_form_equals_definition():
    """The M = 0 closed forms and the channel definitions on G = beta G0 delta agree (two
    independent statements of the same sums)."""
    for nb in (1, 3):
        g = _wl(nb=nb, L=2, n_freq=4, w_ex=(0, 1), channels=ALL_CHANNELS, zeeman=0.29).geometry
        G = bf.g_single_vertex(g, [None, None])
        for ch in g.channels:
            a = bf.g4_free(g, ch, 1.0)
            b = bf.g4_definition(g, G, ch, 1.0)
            assert bf.difference(a, b)["l_inf"] < 1e-14


@pytest.mark.parametrize("variant,channel", [
    ("pp_comment", geo.PARTICLE_PARTICLE_UP_DOWN),
    ("crossed_transposed", geo.PARTICLE_HOLE_MAGNETIC),
    ("crossed_transposed", geo.PARTICLE_HOLE_CHARGE),
    ("transverse_same_spin", geo.PARTICLE_HOLE_TRANSVERSE),
    ("mag_spin_sum", geo.PARTICLE_HOLE_MAGNETIC),
])
@pytest.mark.parametrize("name", ["1band_3x3", "3band_2x2"])
def test_definition_checks_detect_planted_mistakes(oracle_lib, name, variant, channel):
    """Each plausible mistake, planted in the closed-form side on the CPU, is far from the oracle:
    the single-vertex check pins argument order, spin pairing and spin weights."""
    g, verts, meas, acc = _single_vertex_case(oracle_lib, name, 11)
    G = bf.g_single_vertex(g, verts)
    c = g.channels.index(channel)
    got = acc.get_G4()[c]
    right = bf.g4_definition(g, G, channel, meas.sign).ravel(order="F")
    wrong = bf.g4_definition(g, G, channel, meas.sign, variant=variant).ravel(order="F")
    assert bf.difference(right, got)["l_inf"] < 1e-13
    assert bf.difference(wrong, got)["l_inf"] > 1e-3


def test_zeeman_free_forms_see_spin_structure(oracle_lib):
    """With a Zeeman-split G0 and M = 0, the magnetic direct term is non-zero and the s / -s
    pairings differ, so a spin-independent closed form no longer matches the oracle."""
    wl = _wl(nb=3, L=2, n_freq=4, w_ex=(0, 1), channels=ALL_CHANNELS, zeeman=0.37)
    g = wl.geometry
    meas = wl.draw(1)[0]
    meas.M = [np.zeros_like(x) for x in meas.M]
    acc = oracle_lib.OracleAccumulator(g)
    acc.accumulate(meas)
    g_sym = _wl(nb=3, L=2, n_freq=4, w_ex=(0, 1), channels=ALL_CHANNELS).geometry
    g_sym.g0 = np.repeat(g.g0[:, :, :1], 2, axis=2)  # spin slot 0 for both spins
    for c, ch in enumerate(g.channels):
        exact = bf.g4_free(g, ch, meas.sign).ravel(order="F")
        assert bf.difference(exact, acc.get_G4()[c])["l_inf"] < 1e-14
        if ch != geo.PARTICLE_HOLE_CHARGE:
            sym = bf.g4_free(g_sym, ch, meas.sign).ravel(order="F")
            assert bf.difference(sym, acc.get_G4()[c])["l_inf"] > 1e-3, geo.CHANNEL_NAMES[ch]


def test_normalize_g4():
    """G4 / (accumulated sign beta^2) (ctaux_cluster_solver.hpp:523-527), with mixed signs: two
    measurements of sign +1 and one of -1 leave an accumulated sign of 1."""
    rs = synth.Stream(99)
    g4 = rs.normal(64) + 1j * rs.normal(64)
    out = bf.normalize_g4(g4, 4.0, 1.0)
    assert np.allclose(out * 16.0, g4, rtol=1e-15, atol=0)
    out = bf.normalize_g4(g4, 2.0, -3.0)
    assert np.allclose(out * -12.0, g4, rtol=1e-15, atol=0)


def test_compute_m_is_row_scaling():
    """M = diag(exp(V) - 1) N (ctaux_walker.hpp:1889-1916): row i scaled by exp(V_i) - 1, so a
    unit N gives the diagonal and a transposed scaling (columns) differs."""
    rs = synth.Stream(7)
    xv = np.exp(rs.normal(5))
    assert np.array_equal(bf.compute_m(np.eye(5), xv), np.diag(xv - 1.0))
    N = rs.normal(25).reshape(5, 5)
    M = bf.compute_m(N, xv)
    for i in range(5):
        assert np.array_equal(M[i], (xv[i] - 1.0) * N[i])
    assert np.abs(M - N * (xv - 1.0)[None, :]).max() > 1e-3
