"""Parity of the HIP path (dca_amd/libg4acc.so through the C ABI) against the FP64 CPU oracle.

Tolerances (max-normalised l_inf of include/dca/function/util/difference.hpp:71-106):
  FP64 path            1e-12 (G and G4)
  FP32 path (G, G4)    1e-6  (the accuracy gate of the mixed-precision form)
The reference's own CPU-vs-GPU bar is 5e-7 (TT/tp/tp_accumulator_gpu_test.cpp:213).
"""
import ctypes
import math

import numpy as np
import pytest

from dca_amd import geometry as geo
from dca_amd import g4acc, synth
from oracle import bruteforce as bf

pytestmark = pytest.mark.gpu

ALL_CHANNELS = [geo.PARTICLE_HOLE_TRANSVERSE, geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE,
                geo.PARTICLE_PARTICLE_UP_DOWN]

CASES = {
    # name: (nb, L or cluster, n_freq, w_ex, k_ex, n_range)
    "1band_2x2": (1, 2, 4, (0, 1), ((0.0, 0.0), (math.pi, math.pi)), (1, 9)),
    "1band_3x3_shift2": (1, 3, 6, (0, 1, 2), ((0.0, 0.0), (2 * math.pi / 3, 0.0)), (3, 12)),
    "1band_2x2_negative_wex": (1, 2, 4, (-2,), ((math.pi, 0.0),), (1, 8)),
    "3band_2x2": (3, 2, 4, (0, 1), ((0.0, 0.0), (math.pi, math.pi)), (2, 10)),
    "2band_8site": (2, "8site", 4, (0, 1), ((0.0, 0.0), (math.pi, math.pi)), (2, 12)),
    # single band with Nc >= 8: the LDS-tiled G4 kernel
    "1band_4x4_4q": (1, 4, 4, (0, 1, 2), ((0.0, 0.0), (math.pi, 0.0), (math.pi, math.pi),
                                           (math.pi / 2, math.pi / 2)), (10, 40)),
    "1band_8site": (1, "8site", 6, (0, 1), ((0.0, 0.0), (math.pi, math.pi)), (4, 20)),
    "1band_6x6": (1, 6, 2, (0, 1), ((0.0, 0.0), (math.pi, 0.0), (math.pi, math.pi),
                                    (2 * math.pi / 3, 0.0)), (10, 60)),
    # the geometries of the 8x8 and three-band workloads, at reduced frequency counts: the Nc = 64
    # M -> G instantiations and the 1024-thread, four-pairs-per-thread G4 tile; the three-band
    # 8-site tile and its FP32-Kahan sum
    "1band_8x8": (1, 8, 2, (0,), ((0.0, 0.0), (math.pi, math.pi)), (10, 60)),
    "3band_8site": (3, "8site", 4, (0, 1), ((0.0, 0.0), (math.pi, math.pi)), (4, 30)),
    # a Zeeman-split G0: the spin sectors differ, so the magnetic spin-difference term and the
    # s / -s pairings of the transverse and pp channels are exercised
    "1band_4x4_zeeman": (1, 4, 4, (0, 1), ((0.0, 0.0), (math.pi, 0.0)), (8, 30), 0.37),
    "3band_2x2_zeeman": (3, 2, 4, (0, 1), ((0.0, 0.0), (math.pi, math.pi)), (2, 10), 0.37),
}


def case(name, channels=ALL_CHANNELS):
    nb, L, nf, wex, kex, nr = CASES[name][:6]
    zeeman = CASES[name][6] if len(CASES[name]) > 6 else 0.0
    cl = geo.cluster_2d((2, 2), (2, -2)) if L == "8site" else geo.square_cluster(L)
    return synth.tiny_workload(nb=nb, cluster=cl, n_freq=nf, w_ex=wex, k_ex=kex, channels=channels,
                               n_range=nr, seed_cfg=200 + list(CASES).index(name), zeeman=zeeman)


def linf(ref, x):
    return bf.difference(ref, x)["l_inf"]


def oracle_g4(oracle_lib, geom, measurements):
    acc = oracle_lib.OracleAccumulator(geom)
    for m in measurements:
        acc.accumulate(m)
    return acc.get_G4()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_G_matches_oracle(oracle_lib, name, precision, tol):
    wl = case(name)
    acc = g4acc.TpAccumulator(wl.geometry, precision=precision, max_vertices=wl.n_max)
    ora = oracle_lib.OracleAccumulator(wl.geometry)
    for meas in wl.draw(3):
        ora.compute_G(meas)
        ref = ora.get_G()
        got = acc.compute_G(meas)
        err = linf(ref, got)
        print(f"{name} {precision} G l_inf={err:.3e}", flush=True)
        assert err < tol


@pytest.mark.parametrize("name", list(CASES))
def test_G4_kernel_from_oracle_G(oracle_lib, name):
    """The G4 update alone: feed the oracle's G, compare every channel (FP64)."""
    wl = case(name)
    g = wl.geometry
    acc = g4acc.TpAccumulator(g, precision="fp64", max_vertices=wl.n_max)
    acc.resetAccumulation()
    ora = oracle_lib.OracleAccumulator(g)
    for meas in wl.draw(3):
        ora.accumulate(meas)
        acc.accumulate_G(ora.get_G(), meas.sign)
    for ch, ref, got in zip(g.channels, ora.get_G4(), acc.get_G4()):
        err = linf(ref, got)
        print(f"{name} {geo.CHANNEL_NAMES[ch]} l_inf={err:.3e}", flush=True)
        assert err < 1e-13


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision,sum_mode,tol", [("fp64", "fp64", 1e-12), ("fp32", "fp64", 1e-6),
                                                     ("fp32", "fp32_kahan", 1e-6)])
def test_accumulate_matches_oracle(oracle_lib, name, precision, sum_mode, tol):
    wl = case(name)
    g = wl.geometry
    ms = wl.draw(4)
    ref = oracle_g4(oracle_lib, g, ms)
    acc = g4acc.TpAccumulator(g, precision=precision, sum_mode=sum_mode, max_vertices=wl.n_max)
    acc.resetAccumulation()
    for m in ms:
        acc.accumulate(m)
    acc.finalize()
    for ch, r, x in zip(g.channels, ref, acc.get_G4()):
        err = linf(r, x)
        print(f"{name} {precision}/{sum_mode} {geo.CHANNEL_NAMES[ch]} l_inf={err:.3e}", flush=True)
        assert err < tol


@pytest.mark.parametrize("name", ["1band_8x8", "3band_8site", "1band_4x4_zeeman", "3band_2x2_zeeman",
                                  "1band_6x6", "1band_4x4_4q"])
@pytest.mark.parametrize("precision,sum_mode,tol", [("fp64", "fp64", 1e-12), ("fp32", "fp64", 1e-6),
                                                     ("fp32", "fp32_kahan", 1e-6)])
def test_batched_matches_oracle(oracle_lib, name, precision, sum_mode, tol):
    """A batch of 9 walkers in launches of 4 (the batched tile kernels, their XCD-aware schedule,
    the rank-one GEMM where it is on by default) against the oracle's sequential sum."""
    wl = case(name)
    g = wl.geometry
    ms = wl.draw(9)
    ref = oracle_g4(oracle_lib, g, ms)
    acc = g4acc.TpAccumulator(g, precision=precision, sum_mode=sum_mode, max_walkers=4,
                              max_vertices=wl.n_max)
    acc.resetAccumulation()
    acc.accumulate_batch(ms)
    for ch, r, x in zip(g.channels, ref, acc.get_G4()):
        err = linf(r, x)
        print(f"{name} batched {precision}/{sum_mode} {geo.CHANNEL_NAMES[ch]} l_inf={err:.3e}", flush=True)
        assert err < tol


@pytest.mark.parametrize("name", ["1band_4x4_zeeman", "3band_2x2_zeeman", "1band_8x8"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_single_vertex_matches_definition(name, precision, tol):
    """One vertex per spin: G4 of the HIP path against the channel definitions evaluated on the
    closed-form G (oracle/bruteforce.py), independently of the oracle's loops."""
    wl = case(name)
    g = wl.geometry
    rs = synth.Stream(4242)
    verts, n, Ms, taus, sites, bands = [], [], [], [], [], []
    for s in range(2):
        u = rs.uniform(4)
        v = (u[0] * g.beta, int(u[1] * g.nc), int(u[2] * g.nb), g.nc * (1.0 + u[3]))
        verts.append(v)
        n.append(1)
        Ms.append(np.full((1, 1), v[3], order="F"))
        taus.append(np.array([v[0]]))
        sites.append(np.array([v[1]], np.int32))
        bands.append(np.array([v[2]], np.int32))
    meas = synth.Measurement(n=n, M=Ms, tau=taus, site=sites, band=bands, sign=-1.0)
    G = bf.g_single_vertex(g, verts)
    acc = g4acc.TpAccumulator(g, precision=precision, max_vertices=4)
    acc.resetAccumulation()
    acc.accumulate(meas)
    for ch, x in zip(g.channels, acc.get_G4()):
        ref = bf.g4_definition(g, G, ch, meas.sign).ravel(order="F")
        err = linf(ref, x)
        print(f"{name} single vertex {precision} {geo.CHANNEL_NAMES[ch]} l_inf={err:.3e}", flush=True)
        assert err < tol


@pytest.mark.parametrize("name", ["1band_4x4_4q", "3band_2x2"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("memory", ["host", "device"])
def test_queued_accumulate_equals_batch(oracle_lib, name, precision, memory):
    """One accumulate per measurement through the handle's queue (max_walkers = 4) gives the
    same G4 as accumulate_batch over the same chunks; the caller's buffers are free after
    sync_copy (host arrays are overwritten right after the call)."""
    import torch
    from dca_amd.devpool import DevicePool
    wl = case(name)
    g = wl.geometry
    ms = wl.draw(11)
    bat = g4acc.TpAccumulator(g, precision=precision, max_walkers=4, max_vertices=wl.n_max)
    bat.resetAccumulation()
    bat.accumulate_batch(ms)
    que = g4acc.TpAccumulator(g, precision=precision, max_walkers=4, max_vertices=wl.n_max)
    que.resetAccumulation()
    if memory == "host":
        for meas in ms:
            copy = synth.Measurement(n=list(meas.n), M=[m.copy() for m in meas.M],
                                     tau=[t.copy() for t in meas.tau], site=[x.copy() for x in meas.site],
                                     band=[x.copy() for x in meas.band], sign=meas.sign)
            que.accumulate_queued(copy)
            for m in copy.M:
                m[...] = np.nan  # the queue holds its own copy
    else:
        pool = DevicePool(torch, ms)
        for w in pool.walkers:
            que.accumulate_queued(None, memory=g4acc.DEVICE, walker=w)
            que.sync_copy()
    ref = oracle_g4(oracle_lib, g, ms)
    for r, a, b in zip(ref, bat.get_G4(), que.get_G4()):
        assert np.array_equal(a, b)
        assert linf(r, b) < (1e-12 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("name", ["1band_3x3_shift2", "3band_2x2"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batched_equals_sequential(oracle_lib, name, precision):
    wl = case(name)
    g = wl.geometry
    ms = wl.draw(7)
    seq = g4acc.TpAccumulator(g, precision=precision, max_vertices=wl.n_max)
    seq.resetAccumulation()
    for m in ms:
        seq.accumulate(m)
    bat = g4acc.TpAccumulator(g, precision=precision, max_walkers=3, max_vertices=wl.n_max)
    bat.resetAccumulation()
    bat.accumulate_batch(ms)  # 7 walkers through a 3-walker launch: chunks of 3, 3, 1
    for a, b in zip(seq.get_G4(), bat.get_G4()):
        assert linf(a, b) < (1e-13 if precision == "fp64" else 1e-6)
    ref = oracle_g4(oracle_lib, g, ms)
    for r, b in zip(ref, bat.get_G4()):
        assert linf(r, b) < (1e-12 if precision == "fp64" else 1e-6)


def test_sign_linearity_and_empty_configuration():
    wl = case("1band_2x2")
    g = wl.geometry
    m = wl.draw(1)[0]
    acc = g4acc.TpAccumulator(g, max_vertices=wl.n_max)
    acc.resetAccumulation()
    m.sign = 1.0
    acc.accumulate(m)
    plus = acc.get_G4()
    acc.resetAccumulation()
    m.sign = -1.0
    acc.accumulate(m)
    minus = acc.get_G4()
    for p, q in zip(plus, minus):
        assert np.max(np.abs(p + q)) <= 1e-14 * np.abs(p).max()
    empty = synth.Measurement(n=[0, 0], M=[np.zeros((0, 0))] * 2, tau=[np.zeros(0)] * 2,
                              site=[np.zeros(0, np.int32)] * 2, band=[np.zeros(0, np.int32)] * 2,
                              sign=1.0)
    acc.accumulate(empty)
    for a, b in zip(minus, acc.get_G4()):
        assert np.array_equal(a, b)


def test_one_empty_spin_sector(oracle_lib):
    wl = case("3band_2x2")
    g = wl.geometry
    m = wl.draw(1)[0]
    m.n[1] = 0
    m.M[1] = np.zeros((0, 0))
    m.tau[1], m.site[1], m.band[1] = np.zeros(0), np.zeros(0, np.int32), np.zeros(0, np.int32)
    ref = oracle_g4(oracle_lib, g, [m])
    acc = g4acc.TpAccumulator(g, max_vertices=wl.n_max)
    acc.resetAccumulation()
    acc.accumulate(m)
    for r, x in zip(ref, acc.get_G4()):
        assert linf(r, x) < 1e-12


def test_sum_to_and_errors(oracle_lib):
    wl = case("1band_2x2")
    g = wl.geometry
    m1, m2 = wl.draw(2)
    a = g4acc.TpAccumulator(g, max_vertices=wl.n_max)
    b = g4acc.TpAccumulator(g, max_vertices=wl.n_max)
    with pytest.raises(g4acc.LogicError, match="before resetAccumulation"):
        a.accumulate(m1)
    a.resetAccumulation()
    b.resetAccumulation()
    a.accumulate(m1)
    b.accumulate(m2)
    a.sumTo(b)
    ref = oracle_g4(oracle_lib, g, [m1, m2])
    for r, x in zip(ref, b.get_G4()):
        assert linf(r, x) < 1e-12
    with pytest.raises(g4acc.LogicError, match="no G4"):
        a.get_G4()
    big = wl.draw(1)[0]
    big.n[0] = wl.n_max + 1
    with pytest.raises(g4acc.LogicError, match="max_vertices"):
        b.accumulate(big)
    bad = wl.draw(1)[0]
    bad.site[0] = bad.site[0].copy()
    bad.site[0][0] = g.nc  # out of range
    with pytest.raises(g4acc.LogicError, match="out of range"):
        b.accumulate(bad)


@pytest.mark.parametrize("name", ["1band_4x4_4q", "1band_6x6", "1band_8site", "3band_2x2", "2band_8site"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tiled_equals_generic_kernel(name, precision, monkeypatch):
    """The LDS-tiled G4 kernels (single band, multiband) and the generic per-entry kernel agree."""
    wl = case(name, channels=[geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE,
                              geo.PARTICLE_PARTICLE_UP_DOWN])
    g = wl.geometry
    ms = wl.draw(5)
    tiled = g4acc.TpAccumulator(g, precision=precision, max_walkers=5, max_vertices=wl.n_max)
    monkeypatch.setenv("G4ACC_GENERIC_G4", "1")
    generic = g4acc.TpAccumulator(g, precision=precision, max_walkers=5, max_vertices=wl.n_max)
    for acc in (tiled, generic):
        acc.resetAccumulation()
        acc.accumulate_batch(ms)
    for a, b in zip(tiled.get_G4(), generic.get_G4()):
        err = linf(b, a)
        print(f"{name} {precision} tiled vs generic l_inf={err:.3e}", flush=True)
        # FP32: the tiled kernel sums 8 walker terms in FP32 before its FP64 flush, the generic
        # kernel adds every term in FP64.
        assert err < (1e-14 if precision == "fp64" else 5e-7)


@pytest.mark.parametrize("name", ["1band_4x4_4q", "1band_6x6", "1band_8site"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tile_schedules_agree(oracle_lib, name, precision):
    """The G4 tile schedules (natural, super-tiles, diagonal strips of several widths) cover every
    tile once: same G4 up to summation order, and the oracle's."""
    wl = case(name)
    g = wl.geometry
    ms = wl.draw(5)
    ref = oracle_g4(oracle_lib, g, ms)
    out = []
    # (order, strip): natural, super-tiles, diagonal strips of 2 (default), 1 and 5 positions
    for order, strip in ((0, 2), (1, 2), (2, 2), (2, 1), (2, 5)):
        acc = g4acc.TpAccumulator(g, precision=precision, max_walkers=5, max_vertices=wl.n_max,
                                  tuned=False)
        acc.set_tuning(g4_tile_order=order, g4_tile_strip=strip)
        acc.resetAccumulation()
        acc.accumulate_batch(ms)
        out.append(acc.get_G4())
        acc.close()
    for c, r in enumerate(ref):
        for k in range(1, len(out)):
            err = linf(out[0][c], out[k][c])
            print(f"{name} {precision} schedule {k} vs natural l_inf={err:.3e}", flush=True)
            assert err < (1e-14 if precision == "fp64" else 1e-7)
        assert linf(r, out[2][c]) < (1e-12 if precision == "fp64" else 1e-6)


MAG, CHG, PP = (geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE,
                geo.PARTICLE_PARTICLE_UP_DOWN)


@pytest.mark.parametrize("name", ["1band_4x4_4q", "1band_8site", "1band_3x3_shift2"])
@pytest.mark.parametrize("channels", [(MAG, CHG, PP), (MAG,), (CHG, PP), (MAG, CHG)])
def test_mfma_rank_one_matches_oracle(oracle_lib, name, channels):
    """FP32 G4 with the rank-one terms on the matrix cores: against the FP64 oracle (1e-6) and
    against the vector-ALU tile kernel (5e-7). Nine walkers leave a partial group of four and a
    partial FP32 flush of eight."""
    wl = case(name, channels=list(channels))
    g = wl.geometry
    ms = wl.draw(9)
    ref = oracle_g4(oracle_lib, g, ms)
    mf = g4acc.TpAccumulator(g, precision="fp32", max_walkers=9, max_vertices=wl.n_max, tuned=False)
    vec = g4acc.TpAccumulator(g, precision="fp32", max_walkers=9, max_vertices=wl.n_max, tuned=False)
    mf.set_tuning(g4_mfma=1)
    vec.set_tuning(g4_mfma=0, g4_rank1=0)
    for acc in (mf, vec):
        acc.resetAccumulation()
        acc.accumulate_batch(ms)
    for ch, r, a, b in zip(g.channels, ref, mf.get_G4(), vec.get_G4()):
        e_ref, e_vec = linf(r, a), linf(b, a)
        print(f"{name} channel {ch}: mfma vs oracle {e_ref:.3e}, vs vector kernel {e_vec:.3e}",
              flush=True)
        assert e_ref < 1e-6
        assert e_vec < 5e-7


def test_normalized_g4_is_g4_over_beta2_sign():
    """The solvers' normalisation G4 / (beta^2 * accumulated sign) in FP64; a zero sign is refused."""
    wl = case("1band_4x4_4q")
    g = wl.geometry
    ms = wl.draw(5)
    acc = g4acc.TpAccumulator(g, precision="fp64", max_walkers=5, max_vertices=wl.n_max)
    acc.resetAccumulation()
    acc.accumulate_batch(ms)
    sign_sum = sum(m.sign for m in ms)
    raw = acc.get_G4()
    for a, b in zip(raw, acc.get_G4_normalized(sign_sum)):
        np.testing.assert_allclose(b, a / (g.beta ** 2 * sign_sum), rtol=1e-15, atol=0)
    with pytest.raises(g4acc.LogicError):
        acc.get_G4_normalized(0.0)


def test_accumulate_single_is_the_queued_accumulate():
    """g4acc_accumulate_single (the one-measurement drop-in name) batches like
    g4acc_accumulate_queued: seven host measurements through a queue of three equal the batch."""
    wl = case("1band_4x4_4q")
    g = wl.geometry
    ms = wl.draw(7)
    bat = g4acc.TpAccumulator(g, precision="fp32", max_walkers=3, max_vertices=wl.n_max)
    one = g4acc.TpAccumulator(g, precision="fp32", max_walkers=3, max_vertices=wl.n_max)
    bat.resetAccumulation()
    one.resetAccumulation()
    for k in range(0, 7, 3):
        bat.accumulate_batch(ms[k:k + 3])
    keep = g4acc._Keep()
    for meas in ms:
        w = g4acc.walker_struct(meas, keep)
        one._check(one.lib.g4acc_accumulate_single(one._h, ctypes.byref(w), g4acc.HOST))
    for a, b in zip(bat.get_G4(), one.get_G4()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["1band_4x4_4q", "1band_8site", "1band_3x3_shift2", "1band_6x6"])
@pytest.mark.parametrize("channels", [(MAG, CHG, PP), (MAG,), (CHG,), (MAG, PP), (CHG, PP)])
def test_rank_one_gemm_matches_oracle(oracle_lib, name, channels):
    """FP32 G4 with the batch's rank-one MAGNETIC / CHARGE terms from one matrix-core GEMM
    (g4_rank1): against the FP64 oracle (1e-6) and against the tile kernel that forms them per
    walker (5e-7). Nine walkers leave a partial step of four and a partial flush of eight; the
    6x6 cluster runs the 1024-thread tile (two pairs per thread, two momenta per workgroup)."""
    wl = case(name, channels=list(channels))
    g = wl.geometry
    ms = wl.draw(9)
    ref = oracle_g4(oracle_lib, g, ms)
    r1 = g4acc.TpAccumulator(g, precision="fp32", max_walkers=9, max_vertices=wl.n_max, tuned=False)
    vec = g4acc.TpAccumulator(g, precision="fp32", max_walkers=9, max_vertices=wl.n_max, tuned=False)
    r1.set_tuning(g4_rank1=1)
    vec.set_tuning(g4_rank1=0)
    for acc in (r1, vec):
        acc.resetAccumulation()
        acc.accumulate_batch(ms)
        acc.accumulate_batch(ms[:5])  # a second, smaller batch reuses the R planes
    ref5 = oracle_g4(oracle_lib, g, ms[:5])
    for ch, r, r5, a, b in zip(g.channels, ref, ref5, r1.get_G4(), vec.get_G4()):
        e_ref, e_vec = linf(r + r5, a), linf(b, a)
        print(f"{name} channel {ch}: rank-one GEMM vs oracle {e_ref:.3e}, vs tile kernel {e_vec:.3e}",
              flush=True)
        assert e_ref < 1e-6
        assert e_vec < 5e-7


@pytest.mark.parametrize("name", ["1band_2x2", "1band_4x4_4q", "1band_6x6", "1band_8site"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_equals_unfused_mtog(name, precision, monkeypatch):
    """The register-fused single-band M -> G kernels and the four-stage kernels agree."""
    wl = case(name)
    fused = g4acc.TpAccumulator(wl.geometry, precision=precision, max_vertices=wl.n_max)
    monkeypatch.setenv("G4ACC_UNFUSED_MTOG", "1")
    unfused = g4acc.TpAccumulator(wl.geometry, precision=precision, max_vertices=wl.n_max)
    for meas in wl.draw(2):
        a, b = fused.compute_G(meas), unfused.compute_G(meas)
        err = linf(b, a)
        print(f"{name} {precision} fused vs unfused G l_inf={err:.3e}", flush=True)
        assert err < (1e-14 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("name", ["1band_6x6", "1band_8x8"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_large_cluster_mtog_equals_fused(name, precision, monkeypatch):
    """6x6 and 8x8: the matrix-core kernels (default: in FP32 S3, site transform and epilogue in
    one kernel; in FP64 S3 into the E workspace plus the site-transform kernel) against the
    register-fused right kernel (mtog_large = 0), the four-stage kernels and, in FP32, the E
    workspace path (mtog_large = 1), over batches of two walkers (four slots: the chunked
    workspaces) with uneven vertex counts."""
    wl = case(name)
    ms = wl.draw(2)
    large = g4acc.TpAccumulator(wl.geometry, precision=precision, max_walkers=2, max_vertices=wl.n_max)
    reg = g4acc.TpAccumulator(wl.geometry, precision=precision, max_walkers=2, max_vertices=wl.n_max,
                              tuned=False)
    reg.set_tuning(mtog_large=0)
    monkeypatch.setenv("G4ACC_MTOG_E", "1")
    with_e = g4acc.TpAccumulator(wl.geometry, precision=precision, max_walkers=2, max_vertices=wl.n_max,
                                 tuned=False)
    with_e.set_tuning(mtog_large=1)
    monkeypatch.setenv("G4ACC_UNFUSED_MTOG", "1")
    four = g4acc.TpAccumulator(wl.geometry, precision=precision, max_vertices=wl.n_max)
    tol = 1e-14 if precision == "fp64" else 1e-6
    for meas in ms:
        a, b, c, d = large.compute_G(meas), reg.compute_G(meas), four.compute_G(meas), with_e.compute_G(meas)
        print(f"{name} {precision} large vs fused {linf(b, a):.3e}, vs four-stage {linf(c, a):.3e}, "
              f"vs E workspace {linf(d, a):.3e}", flush=True)
        assert linf(b, a) < tol and linf(c, a) < tol and linf(d, a) < tol
    # G4 of both paths against the FP64 G4. In complex64 this two-frequency case cancels strongly
    # in the transverse channel (both paths reach ~1e-6 there), so the bar is that the matrix-core
    # path is no less accurate than the register path it replaces.
    ref = g4acc.TpAccumulator(wl.geometry, precision="fp64", max_walkers=2, max_vertices=wl.n_max)
    # One slot per M -> G pass: the chunk loop over a batch's slots gives the same G4 as one pass
    # (both handles untuned, so the same kernels sum in the same order).
    base = g4acc.TpAccumulator(wl.geometry, precision=precision, max_walkers=2, max_vertices=wl.n_max,
                               tuned=False)
    one = g4acc.TpAccumulator(wl.geometry, precision=precision, max_walkers=2, max_vertices=wl.n_max,
                              tuned=False)
    one.set_tuning(mtog_chunk=1)
    for acc in (large, reg, with_e, ref, base, one):
        acc.resetAccumulation()
        acc.accumulate_batch(ms)
    for x, o in zip(base.get_G4(), one.get_G4()):
        assert linf(x, o) < 1e-13
    for x, y, z, r in zip(large.get_G4(), reg.get_G4(), with_e.get_G4(), ref.get_G4()):
        print(f"{name} {precision} G4 large {linf(r, x):.3e}, fused {linf(r, y):.3e}, "
              f"E workspace {linf(r, z):.3e} vs FP64", flush=True)
        if precision == "fp64":
            assert linf(r, x) < 1e-12 and linf(r, y) < 1e-12 and linf(r, z) < 1e-12
        else:
            assert linf(r, x) <= max(1e-6, 1.25 * linf(r, y)) and linf(r, z) <= max(1e-6, 1.25 * linf(r, y))


@pytest.mark.parametrize("L,n_freq", [(4, 6), (8, 2)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_transpose_symmetric_tiles(oracle_lib, L, n_freq, precision, tol):
    """With one q = 0 and one v = 0 the magnetic, charge and transverse G4 are symmetric under
    (k1, w1) <-> (k2, w2); the tile kernel then computes the tiles w1 <= w2 and writes each mirror
    through LDS. Against the oracle and against the full set of tiles (g4_sym = 0), for single
    measurements and for batches (several walkers per tile)."""
    channels = [geo.PARTICLE_HOLE_TRANSVERSE, geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE]
    wl = synth.tiny_workload(nb=1, L=L, n_freq=n_freq, w_ex=(0,), k_ex=((0.0, 0.0),), channels=channels,
                             n_range=(10, 50), seed_cfg=640 + L)
    g = wl.geometry
    ms = wl.draw(5)
    ref = oracle_g4(oracle_lib, g, ms)
    for walkers in (1, 3):
        sym = g4acc.TpAccumulator(g, precision=precision, max_walkers=walkers, max_vertices=wl.n_max, tuned=False)
        full = g4acc.TpAccumulator(g, precision=precision, max_walkers=walkers, max_vertices=wl.n_max, tuned=False)
        full.set_tuning(g4_sym=0)
        for acc in (sym, full):
            acc.resetAccumulation()
            acc.accumulate_batch(ms)
        for ch, r, x, y in zip(g.channels, ref, sym.get_G4(), full.get_G4()):
            print(f"L={L} {precision} walkers={walkers} {geo.CHANNEL_NAMES[ch]}: symmetric tiles "
                  f"{linf(r, x):.3e}, all tiles {linf(r, y):.3e}, between {linf(y, x):.3e}", flush=True)
            assert linf(r, x) < tol and linf(y, x) < (1e-13 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("name,dense", [("1band_2x2", False), ("1band_2x2", True),
                                        ("1band_4x4_4q", False), ("1band_8site", False)])
def test_mfma_right_mtog_matches_fp64(name, dense, monkeypatch):
    """The matrix-core right M -> G kernel (complex64; separable and dense site DFT) gives the G
    of the FP64 four-stage kernels to complex64 accuracy."""
    wl = case(name)
    if dense:
        monkeypatch.setenv("G4ACC_DENSE_DFT", "1")
    mf = g4acc.TpAccumulator(wl.geometry, precision="fp32", max_vertices=wl.n_max, tuned=False)
    mf.set_tuning(mtog_mfma=1)
    monkeypatch.setenv("G4ACC_UNFUSED_MTOG", "1")
    ref = g4acc.TpAccumulator(wl.geometry, precision="fp64", max_vertices=wl.n_max)
    for meas in wl.draw(3):
        a, b = mf.compute_G(meas), ref.compute_G(meas)
        err = linf(b, a)
        print(f"{name} dense={dense} matrix-core M->G vs FP64 G l_inf={err:.3e}", flush=True)
        assert err < 1e-6


@pytest.mark.parametrize("name", ["1band_2x2", "1band_4x4_4q", "1band_6x6", "3band_2x2"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_separable_site_dft_equals_dense(name, precision, monkeypatch):
    """On square clusters the sites are relabelled x + L y and the right site DFT runs as two
    L-point passes; G equals the dense-DFT path with the caller's labels."""
    wl = case(name)
    sep = g4acc.TpAccumulator(wl.geometry, precision=precision, max_vertices=wl.n_max)
    monkeypatch.setenv("G4ACC_DENSE_DFT", "1")
    dense = g4acc.TpAccumulator(wl.geometry, precision=precision, max_vertices=wl.n_max)
    for meas in wl.draw(2):
        a, b = sep.compute_G(meas), dense.compute_G(meas)
        err = linf(b, a)
        print(f"{name} {precision} separable vs dense site DFT G l_inf={err:.3e}", flush=True)
        assert err < (1e-14 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("name", ["1band_4x4_4q", "3band_2x2"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_equal_eager(name, precision, monkeypatch):
    """Batches replayed from the captured hipGraph give bit-identical G4 to eager launches, on
    every replay, with host synchronisation and eager work between the replays."""
    wl = case(name)
    ms = wl.draw(12)
    graph = g4acc.TpAccumulator(wl.geometry, precision=precision, max_walkers=4, max_vertices=wl.n_max)
    monkeypatch.setenv("G4ACC_GRAPH", "0")
    eager = g4acc.TpAccumulator(wl.geometry, precision=precision, max_walkers=4, max_vertices=wl.n_max)
    graph.resetAccumulation()
    eager.resetAccumulation()
    for i in range(3):
        batch = ms[4 * i:4 * i + 4]
        graph.accumulate_batch(batch)
        eager.accumulate_batch(batch)
        eager.compute_G(batch[0])  # unrelated eager work on the graph's buffers between replays
        graph.compute_G(batch[1])
        for a, b in zip(graph.get_G4(), eager.get_G4()):
            assert np.array_equal(a, b), f"replay {i}"
    graph.accumulate_batch(ms[:3])  # another batch size: a second captured sequence
    eager.accumulate_batch(ms[:3])
    for a, b in zip(graph.get_G4(), eager.get_G4()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["1band_4x4_4q", "3band_2x2", "1band_8x8"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_accumulate_from_n_equals_precomputed_m(oracle_lib, name, precision):
    """Walker hand-off of N and exp(V) (g4acc_accumulate_batch_n): M = diag(exp(V) - 1) N formed
    inside the M -> G step gives the same G4, bit for bit, as accumulating that M computed on the
    CPU (oracle/bruteforce.py compute_m), and matches the oracle."""
    wl = case(name)
    g = wl.geometry
    rs = synth.Stream(31337)
    ms = wl.draw(5)
    pairs = [synth.walker_n_pair(m, rs) for m in ms]
    ms_m = [synth.Measurement(n=list(p.n), M=[bf.compute_m(p.M[s], x[s]) for s in range(2)], tau=p.tau,
                              site=p.site, band=p.band, sign=p.sign) for p, x in pairs]
    a = g4acc.TpAccumulator(g, precision=precision, max_walkers=2, max_vertices=wl.n_max)
    a.resetAccumulation()
    a.accumulate_batch_n([p for p, _ in pairs], [x for _, x in pairs])
    b = g4acc.TpAccumulator(g, precision=precision, max_walkers=2, max_vertices=wl.n_max)
    b.resetAccumulation()
    b.accumulate_batch(ms_m)
    ref = oracle_g4(oracle_lib, g, ms_m)
    tol = 1e-12 if precision == "fp64" else 1e-6
    for r, x, y in zip(ref, a.get_G4(), b.get_G4()):
        assert np.array_equal(x, y)
        assert linf(r, x) < tol


def test_compute_m_on_device():
    """g4acc_compute_m (CtauxWalker::computeM's device branch) equals the CPU restatement bit for
    bit, with leading dimensions larger than n."""
    import torch
    wl = case("1band_4x4_4q")
    acc = g4acc.TpAccumulator(wl.geometry, precision="fp64", max_vertices=wl.n_max)
    rs = synth.Stream(99)
    for n, ldn, ldm in [(1, 1, 1), (37, 40, 37), (300, 300, 305)]:
        N = rs.normal(ldn * n).reshape(ldn, n, order="F")
        xv = np.exp(rs.normal(n) * 0.5)
        dN = torch.from_numpy(np.asfortranarray(N).ravel(order="F")).cuda()
        dx = torch.from_numpy(xv).cuda()
        dM = torch.full((ldm * n,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()  # inputs written on torch's stream; the library runs on its own
        acc.compute_m(n, dx.data_ptr(), dN.data_ptr(), ldn, dM.data_ptr(), ldm)
        acc.synchronizeCopy()
        M = dM.cpu().numpy().reshape(ldm, n, order="F")[:n]
        assert np.array_equal(M, bf.compute_m(N[:n], xv))


@pytest.mark.parametrize("name", ["1band_4x4_4q", "3band_2x2"])
def test_normalized_g4_and_hdf5_export(oracle_lib, tmp_path, name):
    """Mixed-sign measurements: g4acc_get_g4_normalized against the oracle's G4 normalised by the
    restated solver rule (G4 / (accumulated sign beta^2), ctaux_cluster_solver.hpp:523-527), and
    the HDF5 export of the same numbers in the reference's format, read back bit for bit."""
    from dca_amd import g4io
    wl = case(name)
    g = wl.geometry
    ms = wl.draw(6)
    for i, m in enumerate(ms):
        m.sign = -1.0 if i in (1, 4) else 1.0
    acc_sign = sum(m.sign for m in ms)
    ref = [bf.normalize_g4(x, g.beta, acc_sign) for x in oracle_g4(oracle_lib, g, ms)]
    acc = g4acc.TpAccumulator(g, precision="fp64", max_walkers=3, max_vertices=wl.n_max)
    acc.resetAccumulation()
    acc.accumulate_batch(ms)
    got = acc.get_G4_normalized(acc_sign)
    for r, x in zip(ref, got):
        assert linf(r, x) < 1e-12
    path = str(tmp_path / "g4.hdf5")
    acc.write_G4_hdf5(path, accumulated_sign=acc_sign)
    for ch, x in zip(g.channels, got):
        back, dims = g4io.read_function(path, "/functions/G4_" + g4io.channel_name(ch))
        assert dims == g.g4_shape()
        assert np.array_equal(back, x)
