"""Parity of the general-model path on the GPU (g4acc_create_model, g4acc_accumulate_general,
g4acc_compute_G_general) against the FP64 oracle: complex M, complex signs, non-density-density
vertices (columns of M sorted by the vertices' left labels) and models that are not spin
symmetric (pp and magnetic from spin sector 0, tp_accumulator_cpu.hpp:668-757).

Tolerances as in test_gpu_parity.py: max-normalised l_inf 1e-12 (FP64) and 1e-6 (FP32).
"""
import math

import numpy as np
import pytest

from dca_amd import geometry as geo
from dca_amd import g4acc, synth
from oracle import bruteforce as bf

pytestmark = pytest.mark.gpu

CH_ALL = [geo.PARTICLE_HOLE_TRANSVERSE, geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE,
          geo.PARTICLE_PARTICLE_UP_DOWN]
CH_NSS = [geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_PARTICLE_UP_DOWN]

CASES = {
    # name: (nb, L, spin_symmetric, complex_m, non_density_density, n_range)
    "complex_nd_2band": (2, 2, True, True, True, (3, 14)),
    "complex_1band_3x3": (1, 3, True, True, False, (4, 16)),
    "real_nd_1band_4x4": (1, 4, True, False, True, (10, 40)),
    "nss_complex_nd_2band": (2, 2, False, True, True, (3, 14)),
    "nss_real_3band": (3, 2, False, False, False, (3, 12)),
}


def setup(name):
    nb, L, ss, cm, nd, nr = CASES[name]
    g = synth.tiny_workload(nb=nb, L=L, n_freq=4, w_ex=(0, 1), k_ex=((0.0, 0.0), (2 * math.pi / L, 0.0)),
                            channels=CH_ALL if ss else CH_NSS, zeeman=0.23).geometry
    rs = synth.Stream(300 + list(CASES).index(name))
    ms = []
    for i in range(5):
        n = (rs.randint(*nr), rs.randint(*nr))
        ms.append(synth.general_measurement(rs, g.nb, g.nc, g.beta, n, complex_m=cm, non_density=nd,
                                            phase=0.4 + 1.3 * i))
    ms[2].n = [0, ms[2].n[1]]  # an empty spin sector
    ms[2].M[0] = np.zeros((0, 0), ms[2].M[1].dtype)
    for a in ("tau", "site", "band") + (("site_l", "band_l") if nd else ()):
        getattr(ms[2], a)[0] = getattr(ms[2], a)[0][:0]
    return g, ss, cm, nd, ms


def handle(g, precision, ss, cm, nd, walkers=3, sum_mode="fp64", n_max=64):
    acc = g4acc.TpAccumulator(g, precision=precision, sum_mode=sum_mode, max_walkers=walkers, max_vertices=n_max,
                              spin_symmetric=ss, complex_m=cm, non_density_density=nd)
    acc.resetAccumulation()
    return acc


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_general_G_matches_oracle(oracle_lib, name, precision, tol):
    g, ss, cm, nd, ms = setup(name)
    acc = handle(g, precision, ss, cm, nd)
    ora = oracle_lib.OracleAccumulator(g, spin_symmetric=ss)
    for meas in ms[:3]:
        ora.compute_G(meas)
        err = bf.difference(ora.get_G(), acc.compute_G(meas))["l_inf"]
        print(f"{name} {precision} G l_inf={err:.3e}", flush=True)
        assert err < tol


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision,tol,sum_mode", [("fp64", 1e-12, "fp64"), ("fp32", 1e-6, "fp64"),
                                                    ("fp32", 1e-6, "fp32_kahan")])
def test_general_G4_matches_oracle(oracle_lib, name, precision, tol, sum_mode):
    """Five measurements with complex signs (one with an empty spin sector) in batches of three."""
    g, ss, cm, nd, ms = setup(name)
    ora = oracle_lib.OracleAccumulator(g, spin_symmetric=ss)
    for m in ms:
        ora.accumulate(m)
    acc = handle(g, precision, ss, cm, nd, sum_mode=sum_mode)
    acc.accumulate_batch(ms)
    for ch, r, x in zip(g.channels, ora.get_G4(), acc.get_G4()):
        err = bf.difference(r, x)["l_inf"]
        print(f"{name} {precision} {sum_mode} {geo.CHANNEL_NAMES[ch]} l_inf={err:.3e}", flush=True)
        assert err < tol


def test_density_density_walkers_on_general_handle(oracle_lib):
    """Real, density-density measurements (no left labels, sign +-1) through a handle of the
    general model equal the oracle, as through the default handle."""
    wl = synth.tiny_workload(nb=2, L=2, n_freq=4, w_ex=(0, 1), k_ex=((0.0, 0.0), (math.pi, math.pi)),
                             channels=CH_ALL, n_range=(2, 10), seed_cfg=321)
    ms = wl.draw(4)
    ora = oracle_lib.OracleAccumulator(wl.geometry)
    for m in ms:
        ora.accumulate(m)
    acc = handle(wl.geometry, "fp64", True, True, True, walkers=2)
    for m in ms:
        m.M = [x.astype(np.complex128) for x in m.M]
    acc.accumulate_batch(ms)
    for r, x in zip(ora.get_G4(), acc.get_G4()):
        assert bf.difference(r, x)["l_inf"] < 1e-12


def test_nss_handle_refuses_other_channels():
    g = synth.tiny_workload(nb=2, L=2, n_freq=4, channels=[geo.PARTICLE_HOLE_CHARGE]).geometry
    with pytest.raises(g4acc.LogicError, match="four point type"):
        g4acc.TpAccumulator(g, spin_symmetric=False)


def test_complex_m_needs_complex_handle():
    g, ss, cm, nd, ms = setup("complex_1band_3x3")
    acc = handle(g, "fp64", True, False, False)
    with pytest.raises(g4acc.LogicError, match="complex M"):
        acc.accumulate_batch(ms[:1])
