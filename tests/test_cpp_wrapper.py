"""The C++ TpAccumulator mirror (include/g4acc.hpp) driven like the CT-AUX accumulator drives
the reference class, against the FP64 oracle (GPU)."""
import math
import os
import subprocess

import numpy as np
import pytest

from dca_amd import geometry as geo
from dca_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "_build", "tp_accumulator_wrapper_test")


def write_case(path, wl, measurements, expected):
    g = wl.geometry
    with open(path, "wb") as f:
        np.array([g.nb, g.nc, g.n_freq, len(g.w_ex), len(g.k_ex), len(g.channels),
                  len(measurements), g.g4_size()], np.int32).tofile(f)
        np.array([g.beta], np.float64).tofile(f)
        for a in (g.w_ex, g.k_ex, g.channels):
            np.asarray(a, np.int32).tofile(f)
        np.asarray(g.cluster.add.ravel(order="F"), np.int32).tofile(f)
        np.asarray(g.cluster.sub.ravel(order="F"), np.int32).tofile(f)
        np.asarray(g.cluster.dft.ravel(order="F"), np.complex128).tofile(f)
        geo.g0_flat(g.g0).tofile(f)
        for m in measurements:
            np.array(m.n, np.int32).tofile(f)
            np.array([m.sign], np.float64).tofile(f)
            for s in range(2):
                np.asarray(m.M[s], np.float64).ravel(order="F").tofile(f)
                np.asarray(m.tau[s], np.float64).tofile(f)
                np.asarray(m.site[s], np.int32).tofile(f)
                np.asarray(m.band[s], np.int32).tofile(f)
        for e in expected:
            np.asarray(e, np.complex128).tofile(f)


def write_case_general(path, g, model, measurements, expected):
    """The general-model case file of tests/cpp/tp_accumulator_general_test.cpp: the geometry,
    the model flags (spin_symmetric, complex_m, non_density_density), then per measurement n[2],
    the complex sign, and per spin M (complex128 or float64, column-major), tau, site, band and,
    for non-density-density models, the left site and band; then the expected G4 per channel."""
    ss, cm, nd = model
    with open(path, "wb") as f:
        np.array([g.nb, g.nc, g.n_freq, len(g.w_ex), len(g.k_ex), len(g.channels),
                  len(measurements), g.g4_size(), int(ss), int(cm), int(nd)], np.int32).tofile(f)
        np.array([g.beta], np.float64).tofile(f)
        for a in (g.w_ex, g.k_ex, g.channels):
            np.asarray(a, np.int32).tofile(f)
        np.asarray(g.cluster.add.ravel(order="F"), np.int32).tofile(f)
        np.asarray(g.cluster.sub.ravel(order="F"), np.int32).tofile(f)
        np.asarray(g.cluster.dft.ravel(order="F"), np.complex128).tofile(f)
        geo.g0_flat(g.g0).tofile(f)
        for m in measurements:
            np.array(m.n, np.int32).tofile(f)
            z = complex(m.sign)
            np.array([z.real, z.imag], np.float64).tofile(f)
            for s in range(2):
                np.asarray(m.M[s], np.complex128 if cm else np.float64).ravel(order="F").tofile(f)
                np.asarray(m.tau[s], np.float64).tofile(f)
                np.asarray(m.site[s], np.int32).tofile(f)
                np.asarray(m.band[s], np.int32).tofile(f)
                if nd:
                    np.asarray(m.site_l[s], np.int32).tofile(f)
                    np.asarray(m.band_l[s], np.int32).tofile(f)
        for e in expected:
            np.asarray(e, np.complex128).tofile(f)


def test_case_file_layout(tmp_path):
    """The case writer produces exactly what the C++ reader consumes (CPU check of the format)."""
    wl = synth.tiny_workload(nb=1, L=2, n_freq=4, w_ex=(0,), k_ex=((0.0, 0.0),), n_range=(1, 3))
    ms = wl.draw(2)
    g = wl.geometry
    exp = [np.zeros(g.g4_size(), np.complex128) for _ in g.channels]
    p = tmp_path / "case.bin"
    write_case(p, wl, ms, exp)
    n_ext = g.n_ext
    size = 8 * 4 + 8 + 4 * (len(g.w_ex) + len(g.k_ex) + len(g.channels) + 2 * g.nc ** 2) \
        + 16 * (g.nc ** 2 + g.nb ** 2 * 2 * g.nc * n_ext) \
        + sum(4 * 2 + 8 + sum(8 * n * n + 8 * n + 8 * n for n in m.n) for m in ms) \
        + 16 * g.g4_size() * len(g.channels)
    assert p.stat().st_size == size


def test_mirror_host_logic():
    """The mirror's std::logic_error cases that precede any device call (CPU): non-density-density
    vertices and too few single-particle frequencies (tests/cpp/host_logic_test.cpp)."""
    exe = os.path.join(HERE, "cpp", "_build", "host_logic_test")
    if not os.path.exists(exe):
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(HERE, "..", "include"),
                        os.path.join(HERE, "cpp", "host_logic_test.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "OK", out.stdout


def test_site_fft_host():
    """The register site transforms of the M -> G kernels (dca_amd/csrc/site_fft.hpp), run on
    the CPU against the direct double sum (tests/cpp/site_fft_test.cpp)."""
    exe = os.path.join(HERE, "cpp", "_build", "site_fft_test")
    if not os.path.exists(exe):
        os.makedirs(os.path.dirname(exe), exist_ok=True)
        subprocess.run(["g++", "-std=c++17", "-O2", "-Wno-unknown-pragmas", "-D__HIP_PLATFORM_AMD__",
                        "-I/opt/rocm/include", os.path.join(HERE, "cpp", "site_fft_test.cpp"), "-o", exe],
                       check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "OK", out.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("nb,L,precision,tol", [(1, 4, 0, 1e-12), (1, 4, 1, 1e-6), (3, 2, 0, 1e-12)])
def test_cpp_mirror_matches_oracle(oracle_lib, tmp_path, nb, L, precision, tol):
    assert os.path.exists(BIN), "build the driver with tests/cpp/build.sh (__graft_entry__.build())"
    k_ex = ((0.0, 0.0), (math.pi, math.pi))
    wl = synth.tiny_workload(nb=nb, L=L, n_freq=4, w_ex=(0, 1), k_ex=k_ex, n_range=(2, 16),
                             seed_cfg=300 + nb)
    ms = wl.draw(4)
    ref = oracle_lib.OracleAccumulator(wl.geometry)
    for m in ms:
        ref.accumulate(m)
    p = tmp_path / "case.bin"
    write_case(p, wl, ms, ref.get_G4())
    print(f"C++ mirror nb={nb} L={L} precision={precision}", flush=True)
    r = subprocess.run([BIN, str(p), str(precision), str(tol)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "OK" in r.stdout


GENERAL_BIN = os.path.join(HERE, "cpp", "_build", "tp_accumulator_general_test")


@pytest.mark.gpu
@pytest.mark.parametrize("nb,L,model", [(2, 2, (1, 1, 1)), (1, 3, (1, 1, 0)), (2, 2, (0, 1, 1)),
                                        (3, 2, (0, 0, 0))])
@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-6)])
def test_cpp_mirror_general_model(oracle_lib, tmp_path, nb, L, model, precision, tol):
    """The C++ mirror on the general model (complex M and signs, left labels apart from the right
    ones, not spin symmetric) through g4acc_accumulate_general, against the FP64 oracle."""
    assert os.path.exists(GENERAL_BIN), "build the driver with tests/cpp/build.sh (__graft_entry__.build())"
    ss, cm, nd = model
    channels = ([geo.PARTICLE_HOLE_TRANSVERSE, geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE,
                 geo.PARTICLE_PARTICLE_UP_DOWN] if ss else
                [geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_PARTICLE_UP_DOWN])
    g = synth.tiny_workload(nb=nb, L=L, n_freq=4, w_ex=(0, 1), k_ex=((0.0, 0.0), (2 * math.pi / L, 0.0)),
                            channels=channels, zeeman=0.23).geometry
    rs = synth.Stream(700 + 10 * nb + L + 3 * ss + 5 * cm + 7 * nd)
    ms = [synth.general_measurement(rs, g.nb, g.nc, g.beta, (rs.randint(3, 14), rs.randint(3, 14)),
                                    complex_m=bool(cm), non_density=bool(nd), phase=0.4 + 1.3 * i)
          for i in range(4)]
    ref = oracle_lib.OracleAccumulator(g, spin_symmetric=bool(ss))
    for m in ms:
        ref.accumulate(m)
    p = tmp_path / "case.bin"
    write_case_general(p, g, model, ms, ref.get_G4())
    print(f"C++ mirror general nb={nb} L={L} model={model} precision={precision}", flush=True)
    r = subprocess.run([GENERAL_BIN, str(p), str(precision), str(tol)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "OK" in r.stdout
