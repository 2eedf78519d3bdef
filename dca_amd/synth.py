"""Synthetic, portable, seeded measurement streams for the two-particle accumulator.

Every measurement is what a CT-AUX/CT-INT walker hands to TpAccumulator::accumulate
(tp_accumulator_cpu.hpp:231-252): per spin sector s a real N_s x N_s matrix M_s (column-major),
the vertex configuration (tau, site, band) and a sign.

The stream depends only on the raw 64-bit output of numpy's PCG64 bit generator (stable across
numpy versions); distributions are this module's own: uniform u = (x >> 11) * 2^-53, normal by
Box-Muller, Poisson by inversion.

The five workloads are the accumulator configurations this project measures:
  1  1 band, 2x2 cluster, beta 2, 16 frequencies, w_ex {0}, k_ex {(0,0)}, magnetic, N = 32
  2  1 band, 4x4 cluster, beta 10, 64 frequencies, w_ex {0,1,2}, 4 k_ex, three channels,
     N ~ Poisson(128) clipped to [64, 256], 256 walkers per batch
  3  1 band, 6x6 cluster, beta 20, 128 frequencies, w_ex {0,1}, 4 k_ex, magnetic + pp,
     N ~ U[50, 600], 64 walkers per batch
  4  3 bands, 8-site cluster (2,2),(2,-2), beta 15, 64 frequencies, w_ex {0,1}, 2 k_ex,
     three channels, N ~ U[100, 400], 32 walkers per batch
  5  1 band, 8x8 cluster, beta 40, 256 frequencies, w_ex {0}, k_ex {(0,0)}, charge,
     N ~ U[800, 2000], 16 walkers per batch
M entries are N(0, 1/N); times are uniform in [0, beta), sites and bands uniform; the sign is +1
(workload 1) or +1 with probability 0.9, else -1.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable

import numpy as np

from . import geometry as geo


class Stream:
    """Portable random stream over PCG64 raw output."""

    def __init__(self, seed: int):
        self._bg = np.random.PCG64(seed)

    def raw(self, n: int) -> np.ndarray:
        return self._bg.random_raw(n)

    def uniform(self, n: int) -> np.ndarray:
        return (self.raw(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u = self.uniform(2 * m)
        rad = np.sqrt(-2.0 * np.log1p(-u[:m]))
        ang = 2.0 * math.pi * u[m:]
        z = np.empty(2 * m)
        z[0::2] = rad * np.cos(ang)
        z[1::2] = rad * np.sin(ang)
        return z[:n]

    def poisson(self, lam: float) -> int:
        u = self.uniform(1)[0]
        k, p = 0, math.exp(-lam)
        cdf = p
        while u > cdf and k < 100000:
            k += 1
            p *= lam / k
            cdf += p
        return k

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform(1)[0] * (hi - lo + 1))


@dataclasses.dataclass
class Measurement:
    n: list[int]
    M: list[np.ndarray]      # Fortran-ordered (N, N) float64 (complex128 for complex models)
    tau: list[np.ndarray]    # float64
    site: list[np.ndarray]   # int32, the vertices' right sites (both sides for density-density)
    band: list[np.ndarray]   # int32, the vertices' right bands
    sign: float              # +-1, or a complex phase for complex models
    # Non-density-density vertices: their left sites and bands (None: the same as the right ones).
    site_l: list[np.ndarray] | None = None
    band_l: list[np.ndarray] | None = None


def draw_measurement(rs: Stream, nb: int, nc: int, beta: float, draw_n: Callable[[Stream], int],
                     signed: bool) -> Measurement:
    n, Ms, taus, sites, bands = [], [], [], [], []
    for _ in range(2):
        N = int(draw_n(rs))
        u = rs.uniform(3 * N)
        taus.append(u[:N] * beta)
        sites.append(np.minimum((u[N:2 * N] * nc).astype(np.int32), nc - 1))
        bands.append(np.minimum((u[2 * N:] * nb).astype(np.int32), nb - 1))
        z = rs.normal(N * N) * math.sqrt(1.0 / N) if N else np.zeros(0)
        Ms.append(np.asfortranarray(z.reshape(N, N, order="F")))
        n.append(N)
    sign = 1.0
    if signed:
        sign = 1.0 if rs.uniform(1)[0] < 0.9 else -1.0
    return Measurement(n=n, M=Ms, tau=taus, site=sites, band=bands, sign=sign)


@dataclasses.dataclass
class Workload:
    cfg: int
    geometry: geo.Geometry
    draw_n: Callable[[Stream], int]
    signed: bool
    walkers: int          # walkers per batched launch
    n_max: int            # largest N the distribution can draw
    measurements: int     # measurements of the full workload
    description: str

    def stream(self, seed_offset: int = 0) -> Stream:
        return Stream(0x0DCA0000 + self.cfg + (seed_offset << 8))

    def draw(self, count: int, seed_offset: int = 0) -> list[Measurement]:
        rs = self.stream(seed_offset)
        g = self.geometry
        return [draw_measurement(rs, g.nb, g.nc, g.beta, self.draw_n, self.signed)
                for _ in range(count)]

    def draw_chunks(self, count: int, chunk: int, seed_offset: int = 0):
        """The same stream as draw(count), yielded in lists of at most `chunk` measurements (so a
        long run never holds all of its M matrices at once)."""
        rs = self.stream(seed_offset)
        g = self.geometry
        done = 0
        while done < count:
            k = min(chunk, count - done)
            yield [draw_measurement(rs, g.nb, g.nc, g.beta, self.draw_n, self.signed) for _ in range(k)]
            done += k


def _poisson_clipped(lam, lo, hi):
    return lambda rs: min(max(rs.poisson(lam), lo), hi)


def _uniform_n(lo, hi):
    return lambda rs: rs.randint(lo, hi)


def workload(cfg: int) -> Workload:
    pi = math.pi
    if cfg == 1:
        cl = geo.square_cluster(2)
        g = geo.make_geometry(1, cl, 2.0, 16, [0], [cl.k_index((0, 0))],
                              [geo.PARTICLE_HOLE_MAGNETIC])
        return Workload(1, g, lambda rs: 32, False, 1, 32, 1000,
                        "1 band, 2x2 cluster, beta=2, 16 freqs, 1 k_ex, 1 w_ex, ph-magnetic, N=32")
    if cfg == 2:
        cl = geo.square_cluster(4)
        kex = [cl.k_index(k) for k in [(0, 0), (pi, 0), (pi, pi), (pi / 2, pi / 2)]]
        g = geo.make_geometry(1, cl, 10.0, 64, geo.exchange_frequencies(2, True), kex,
                              [geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE,
                               geo.PARTICLE_PARTICLE_UP_DOWN])
        return Workload(2, g, _poisson_clipped(128, 64, 256), True, 256, 256, 10000,
                        "1 band, 4x4 cluster, beta=10, 64 freqs, 4 k_ex, 3 w_ex, 3 channels, "
                        "N~Poisson(128) in [64,256]")
    if cfg == 3:
        cl = geo.square_cluster(6)
        kex = [cl.k_index(k) for k in [(0, 0), (pi, 0), (pi, pi), (2 * pi / 3, 0)]]
        g = geo.make_geometry(1, cl, 20.0, 128, geo.exchange_frequencies(1, True), kex,
                              [geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_PARTICLE_UP_DOWN])
        return Workload(3, g, _uniform_n(50, 600), True, 64, 600, 64 * 16,
                        "1 band, 6x6 cluster, beta=20, 128 freqs, 4 k_ex, 2 w_ex, magnetic+pp, "
                        "N~U[50,600], 64 walkers")
    if cfg == 4:
        cl = geo.cluster_2d((2, 2), (2, -2))
        kex = [cl.k_index(k) for k in [(0, 0), (pi, pi)]]
        g = geo.make_geometry(3, cl, 15.0, 64, geo.exchange_frequencies(1, True), kex,
                              [geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE,
                               geo.PARTICLE_PARTICLE_UP_DOWN])
        return Workload(4, g, _uniform_n(100, 400), True, 32, 400, 32 * 16,
                        "3 bands, 8-site cluster, beta=15, 64 freqs, 2 k_ex, 2 w_ex, 3 channels, "
                        "N~U[100,400], 32 walkers")
    if cfg == 5:
        cl = geo.square_cluster(8)
        g = geo.make_geometry(1, cl, 40.0, 256, [0], [cl.k_index((0, 0))],
                              [geo.PARTICLE_HOLE_CHARGE])
        return Workload(5, g, _uniform_n(800, 2000), True, 16, 2000, 16 * 8,
                        "1 band, 8x8 cluster, beta=40, 256 freqs, 1 k_ex, 1 w_ex, ph-charge, "
                        "N~U[800,2000], 16 walkers")
    raise ValueError(f"unknown workload {cfg}")


def tiny_workload(nb: int = 1, L: int = 2, n_freq: int = 4, beta: float = 2.0,
                  w_ex=(0,), k_ex=((0.0, 0.0),), channels=None, n_range=(1, 6),
                  signed: bool = True, cluster: geo.Cluster | None = None,
                  seed_cfg: int = 100, zeeman: float = 0.0) -> Workload:
    """Small workloads for parity tests (sizes the oracle finishes in well under a second)."""
    cl = cluster if cluster is not None else geo.square_cluster(L)
    if channels is None:
        channels = [geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE,
                    geo.PARTICLE_PARTICLE_UP_DOWN]
    kex = [cl.k_index(k) for k in k_ex]
    g = geo.make_geometry(nb, cl, beta, n_freq, list(w_ex), kex, channels, zeeman=zeeman)
    lo, hi = n_range
    return Workload(seed_cfg, g, _uniform_n(lo, hi), signed, 1, hi, 4, "tiny")


def walker_n_pair(meas: Measurement, rs: Stream, gamma: float = 0.6):
    """CT-AUX-style hand-off for a measurement: per spin the vertices' exp(V) = exp(-gamma s_HS)
    with a random auxiliary spin s_HS = +-1 (CV::exp_V, ctaux/structs/cv.hpp:13-15 with one
    interaction strength), and the N matrix with M = diag(exp(V) - 1) N, i.e. N = M scaled back.
    Returns (measurement with N in place of M, [exp_v spin 0, exp_v spin 1])."""
    Ns, xvs = [], []
    for s in range(2):
        n = meas.n[s]
        hs = np.where(rs.uniform(n) < 0.5, 1.0, -1.0)
        xv = np.exp(-gamma * hs)
        Ns.append(np.asfortranarray(meas.M[s] / (xv - 1.0)[:, None]) if n else meas.M[s])
        xvs.append(xv)
    return Measurement(n=list(meas.n), M=Ns, tau=meas.tau, site=meas.site, band=meas.band,
                       sign=meas.sign), xvs


def general_measurement(rs: Stream, nb: int, nc: int, beta: float, n: tuple[int, int],
                        complex_m: bool = True, non_density: bool = True,
                        phase: float | None = 0.7) -> Measurement:
    """A measurement of the general model: M with complex N(0, 1/N) entries (real and imaginary
    parts each N(0, 1/2N)), vertices whose left (site, band) are drawn apart from their right ones
    (non-density-density vertices, as CT-INT's SectorEntry holds them), and the sign a complex
    phase exp(i phase) (SignType of a complex model; phase None keeps +1)."""
    Ms, taus, sites, bands, sites_l, bands_l = [], [], [], [], [], []
    for s in range(2):
        N = int(n[s])
        u = rs.uniform(5 * N)
        taus.append(u[:N] * beta)
        sites.append(np.minimum((u[N:2 * N] * nc).astype(np.int32), nc - 1))
        bands.append(np.minimum((u[2 * N:3 * N] * nb).astype(np.int32), nb - 1))
        if non_density:
            sites_l.append(np.minimum((u[3 * N:4 * N] * nc).astype(np.int32), nc - 1))
            bands_l.append(np.minimum((u[4 * N:] * nb).astype(np.int32), nb - 1))
        z = rs.normal(N * N) * math.sqrt(1.0 / N) if N else np.zeros(0)
        if complex_m and N:
            z = (z + 1j * rs.normal(N * N) * math.sqrt(1.0 / N)) * math.sqrt(0.5)
        Ms.append(np.asfortranarray(z.reshape(N, N, order="F")))
    sign = 1.0 if phase is None else complex(math.cos(phase), math.sin(phase))
    return Measurement(n=[int(x) for x in n], M=Ms, tau=taus, site=sites, band=bands, sign=sign,
                       site_l=sites_l if non_density else None, band_l=bands_l if non_density else None)
