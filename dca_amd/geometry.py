"""Cluster and frequency geometry of the two-particle accumulator.

These are the pieces of the reference's domains that fix the shape of G and G4 and their index
arithmetic, restated for explicit arguments instead of static singletons:

* compact / extended vertex frequencies:
  include/dca/phys/domains/time_and_frequency/vertex_frequency_domain.hpp:104-128 (COMPACT)
  and :157-185 (EXTENDED); the extension is max|w_ex| from
  src/phys/domains/time_and_frequency/frequency_exchange_domain.cpp:26-48.
* momentum add / subtract tables:
  include/dca/phys/domains/cluster/cluster_domain_initializer.hpp:688-770
  (add(i, j) = index(k_i + k_j), subtract(i, j) = index(k_j - k_i), both folded into the BZ).
* the cluster DFT matrix T(k, r) = exp(i k.r):
  include/dca/math/function_transform/special_transforms/space_transform_2D.hpp:93-114.
* channel ids and order: include/dca/phys/four_point_type.hpp:27-33; the channel list is
  sorted and made unique (include/dca/phys/parameters/four_point_parameters.hpp:194-196).

A cluster is given by its two integer superlattice vectors (a1, a2) on the square lattice.
Sites are the lattice points of one supercell, k-points the reciprocal superlattice points of the
first Brillouin zone (both reduced to a canonical representative, so the index order is this
module's own, not the reference's).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Sequence

import numpy as np

# FourPointType (include/dca/phys/four_point_type.hpp:27-33)
PARTICLE_HOLE_TRANSVERSE = 1
PARTICLE_HOLE_MAGNETIC = 2
PARTICLE_HOLE_CHARGE = 3
PARTICLE_PARTICLE_UP_DOWN = 4

CHANNEL_NAMES = {
    PARTICLE_HOLE_TRANSVERSE: "PARTICLE_HOLE_TRANSVERSE",
    PARTICLE_HOLE_MAGNETIC: "PARTICLE_HOLE_MAGNETIC",
    PARTICLE_HOLE_CHARGE: "PARTICLE_HOLE_CHARGE",
    PARTICLE_PARTICLE_UP_DOWN: "PARTICLE_PARTICLE_UP_DOWN",
}


def sort_channels(channels: Sequence[int]) -> list[int]:
    """Sorted, duplicate-free channel list (four_point_parameters.hpp:194-196)."""
    return sorted(set(int(c) for c in channels))


def exchange_frequencies(frequency_transfer: int, compute_all_transfers: bool) -> list[int]:
    """FrequencyExchangeDomain elements (frequency_exchange_domain.cpp:26-48)."""
    if compute_all_transfers:
        if frequency_transfer < 0:
            raise ValueError("The frequency transfer for multiple transfers must be non-negative.")
        return list(range(frequency_transfer + 1))
    return [int(frequency_transfer)]


def extension_size(w_ex: Sequence[int]) -> int:
    return max(abs(int(w)) for w in w_ex)


def vertex_frequencies(n_freq: int, beta: float) -> np.ndarray:
    """Matsubara frequencies of a vertex domain of n_freq (even) entries:
    w[n/2 + l] = (2l+1) pi / beta, w[n/2 - 1 - l] = -(2l+1) pi / beta."""
    if n_freq % 2:
        raise ValueError("vertex frequency domains have an even size")
    w = np.empty(n_freq)
    for l in range(n_freq // 2):
        w[n_freq // 2 + l] = math.pi / beta * (1 + 2 * l)
        w[n_freq // 2 - 1 - l] = -math.pi / beta * (1 + 2 * l)
    return w


@dataclasses.dataclass
class Cluster:
    r: np.ndarray      # (nc, 2) site vectors
    k: np.ndarray      # (nc, 2) momentum vectors
    add: np.ndarray    # (nc, nc) int, add[i, j] = index(k_i + k_j)
    sub: np.ndarray    # (nc, nc) int, sub[i, j] = index(k_j - k_i)
    dft: np.ndarray    # (nc, nc) complex, dft[k, r] = exp(i k.r)
    rsub: np.ndarray | None = None  # (nc, nc) int, rsub[a, b] = index(r_b - r_a), sites folded

    @property
    def nc(self) -> int:
        return len(self.r)

    def k_index(self, kvec) -> int:
        return _index_of(self.k, _fold_k(np.asarray(kvec, float)))


def _fold_k(k: np.ndarray) -> np.ndarray:
    """Reduce a momentum modulo the square reciprocal lattice 2*pi*Z^2 into [-pi, pi)."""
    return (k + math.pi) % (2 * math.pi) - math.pi


def _index_of(table: np.ndarray, v: np.ndarray) -> int:
    d = np.abs(table - v[None, :])
    d = np.minimum(d, 2 * math.pi - d)  # periodic distance for momenta
    hit = np.nonzero(np.all(d < 1e-8, axis=1))[0]
    if len(hit) != 1:
        raise ValueError(f"vector {v} not found in cluster table")
    return int(hit[0])


def cluster_2d(a1: Sequence[int], a2: Sequence[int]) -> Cluster:
    """Cluster of the square lattice spanned by integer superlattice vectors a1, a2."""
    A = np.array([a1, a2], dtype=float).T  # columns are the superlattice vectors
    nc = int(round(abs(np.linalg.det(A))))
    Ainv = np.linalg.inv(A)
    # Sites: integer points reduced into the supercell by fractional coordinates.
    span = int(np.abs(A).sum()) + 1
    seen = {}
    for x in range(-span, span + 1):
        for y in range(-span, span + 1):
            f = Ainv @ np.array([x, y], float)
            f = f - np.floor(f + 1e-9)
            key = (round(f[0], 9) % 1.0, round(f[1], 9) % 1.0)
            if key not in seen:
                seen[key] = A @ np.array(key)
    if len(seen) != nc:
        raise ValueError("site enumeration failed")
    r = np.array(sorted(seen.values(), key=lambda v: (round(v[1], 6), round(v[0], 6))))
    r = np.round(r).astype(float)
    # Momenta: reciprocal superlattice points B = 2 pi A^{-T}, folded into [-pi, pi)^2.
    B = 2 * math.pi * Ainv.T
    ks = []
    for m in range(-nc, nc + 1):
        for n in range(-nc, nc + 1):
            kv = _fold_k(B @ np.array([m, n], float))
            if not any(np.all(np.abs(np.minimum(np.abs(kv - q), 2 * math.pi - np.abs(kv - q))) < 1e-8)
                       for q in ks):
                ks.append(kv)
    if len(ks) != nc:
        raise ValueError("k enumeration failed")
    k = np.array(sorted(ks, key=lambda v: (round(v[1], 6), round(v[0], 6))))
    k[np.abs(k) < 1e-12] = 0.0
    add = np.empty((nc, nc), np.int32)
    sub = np.empty((nc, nc), np.int32)
    for i in range(nc):
        for j in range(nc):
            add[i, j] = _index_of(k, _fold_k(k[i] + k[j]))
            sub[i, j] = _index_of(k, _fold_k(k[j] - k[i]))
    dft = np.exp(1j * (k @ r.T))
    # Site differences folded into the supercell (the cluster domain's subtract for sites,
    # cluster_domain_initializer.hpp:688-770): rsub[a, b] = index(r_b - r_a).
    def site_index(v):
        f = Ainv @ v
        f = f - np.floor(f + 1e-9)
        rep = np.round(A @ f)
        hit = np.nonzero(np.all(np.abs(r - rep[None, :]) < 1e-6, axis=1))[0]
        if len(hit) != 1:
            raise ValueError(f"site {v} not found in the cluster")
        return int(hit[0])
    rsub = np.empty((nc, nc), np.int32)
    for a in range(nc):
        for b in range(nc):
            rsub[a, b] = site_index(r[b] - r[a])
    return Cluster(r=r, k=k, add=add, sub=sub, dft=dft, rsub=rsub)


def square_cluster(L: int) -> Cluster:
    return cluster_2d((L, 0), (0, L))


def dispersion_single_band(k: np.ndarray, t: float = 1.0) -> np.ndarray:
    """eps_k = -2 t (cos kx + cos ky)."""
    return -2.0 * t * (np.cos(k[:, 0]) + np.cos(k[:, 1]))


def multiband_hamiltonian(k: np.ndarray, nb: int = 3, seed: int = 0x0DCA0004) -> np.ndarray:
    """H(k) = diag(e) + 2 t (cos kx + cos ky) with a fixed seeded Hermitian nb x nb t and
    on-site energies e = (0, 1, -1) for three bands (evenly spaced in [-1, 1] otherwise).
    Returns (nc, nb, nb) complex."""
    rng = np.random.Generator(np.random.PCG64(seed + nb - 3))
    raw = rng.bit_generator.random_raw(2 * nb * nb)
    u = (raw >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    t = (u[:nb * nb] - 0.5).reshape(nb, nb) + 1j * (u[nb * nb:] - 0.5).reshape(nb, nb)
    t = 0.5 * (t + t.conj().T) * 0.5
    e = [0.0, 1.0, -1.0] if nb == 3 else list(np.linspace(-1.0, 1.0, nb))
    base = np.diag(e).astype(complex)
    c = np.cos(k[:, 0]) + np.cos(k[:, 1])
    return base[None] + 2.0 * c[:, None, None] * t[None]


def free_g0(cluster: Cluster, nb: int, beta: float, n_ext: int, zeeman: float = 0.0) -> np.ndarray:
    """Non-interacting G0 over the EXTENDED vertex frequencies.

    Returns complex128 of shape (nb, nb, 2, nc, nwE) in C order (so that the Fortran-order
    flattening of its transpose matches the (b1, b2, s, k, wE) first-index-fastest layout; use
    g0_flat() for that). Single band: g_s = 1/(i w - eps_k - h_s); several bands:
    (i w - H(k) - h_s)^-1, with h_s = +zeeman for spin slot 0 and -zeeman for slot 1. zeeman = 0
    (the measured workloads) gives a spin-independent G0; a non-zero field makes the two spin
    sectors differ, so the spin-difference term of the magnetic channel and the s / -s pairing of
    the transverse and pp channels no longer vanish or coincide.
    """
    w = vertex_frequencies(n_ext, beta)
    nc = cluster.nc
    out = np.zeros((nb, nb, 2, nc, n_ext), complex)
    h = (float(zeeman), -float(zeeman))
    if nb == 1:
        eps = dispersion_single_band(cluster.k)
        for s in range(2):
            out[0, 0, s] = 1.0 / (1j * w[None, :] - eps[:, None] - h[s])
        return out
    H = multiband_hamiltonian(cluster.k, nb)
    for s in range(2):
        for ik in range(nc):
            for iw in range(n_ext):
                out[:, :, s, ik, iw] = np.linalg.inv((1j * w[iw] - h[s]) * np.eye(nb) - H[ik])
    return out


def g0_flat(g0: np.ndarray) -> np.ndarray:
    """(nb, nb, 2, nc, nwE) array -> flat complex128 vector in (b1, b2, s, k, wE) first-fastest order."""
    return np.ascontiguousarray(g0.ravel(order="F"))


@dataclasses.dataclass
class Geometry:
    """All inputs that fix the accumulator's shape (the reference takes them from Parameters)."""
    nb: int
    cluster: Cluster
    beta: float
    n_freq: int                 # compact vertex frequencies (2 * four-point-fermionic-frequencies)
    w_ex: list[int]
    k_ex: list[int]
    channels: list[int]
    g0: np.ndarray              # (nb, nb, 2, nc, nwE) complex128
    band_phase: np.ndarray | None = None  # (nb, nc) complex, exp(i k.a_b); None = no displacement

    @property
    def nc(self) -> int:
        return self.cluster.nc

    @property
    def ext(self) -> int:
        return extension_size(self.w_ex)

    @property
    def n_ext(self) -> int:
        return self.n_freq + 2 * self.ext

    @property
    def D(self) -> int:
        return self.nb * self.nc * self.n_ext

    def g4_shape(self) -> tuple:
        """G4 dims, first index fastest: (b1, b2, b3, b4, k1, w1, k2, w2, k_ex, w_ex)."""
        nb, nc, nw = self.nb, self.nc, self.n_freq
        return (nb, nb, nb, nb, nc, nw, nc, nw, len(self.k_ex), len(self.w_ex))

    def g4_size(self) -> int:
        return int(np.prod(self.g4_shape()))


def make_geometry(nb: int, cluster: Cluster, beta: float, n_freq: int, w_ex: Sequence[int],
                  k_ex: Sequence[int], channels: Sequence[int],
                  g0: np.ndarray | None = None, zeeman: float = 0.0) -> Geometry:
    w_ex = [int(w) for w in w_ex]
    ext = extension_size(w_ex)
    if g0 is None:
        g0 = free_g0(cluster, nb, beta, n_freq + 2 * ext, zeeman)
    return Geometry(nb=nb, cluster=cluster, beta=float(beta), n_freq=int(n_freq), w_ex=w_ex,
                    k_ex=[int(k) for k in k_ex], channels=sort_channels(channels), g0=g0)
