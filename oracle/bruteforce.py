"""Direct-definition evaluators for tiny inputs (TEST INFRASTRUCTURE ONLY).

These do not follow the reference's algorithm (no sorting, no per-orbital GEMMs, no scatter):
they evaluate the defining sums, so they pin the oracle's NDFT/DFT/G steps independently.

* NDFT definition used by the reference's own tests
  (test/unit/.../accumulation/single_sector_accumulation_test.hpp:226-254,
   test/unit/.../accumulation/tp/ndft/cached_ndft_test.hpp:120-143):
      f(b1, b2, r1, r2, w1, w2) = sum_{i in (b1,r1), j in (b2,r2)} M_ij exp(i (t_i w1 - t_j w2))
* Cluster DFT (space_transform_2D.hpp:41-44): f(k1, k2) = sum_{r1,r2} exp(i(k1 r1 - k2 r2)) f(r1,r2) / Nc
* G = -G0 M G0 + beta G0 delta (tp_accumulator_cpu.hpp:308-375) as a plain band-matrix product.
* Closed forms of G4 for M = 0 (G = beta G0 delta) and a spin-independent G0.
"""
from __future__ import annotations

import numpy as np

from dca_amd import geometry as geo


def ndft_direct(geom: geo.Geometry, meas, spin: int) -> np.ndarray:
    """Returns f[r1, r2, b1, b2, w1E, w2E] by the defining double sum."""
    nb, nc, nE = geom.nb, geom.nc, geom.n_ext
    w = geo.vertex_frequencies(nE, geom.beta)
    n = meas.n[spin]
    out = np.zeros((nc, nc, nb, nb, nE, nE), complex)
    if n == 0:
        return out
    tau = meas.tau[spin]
    ph = np.exp(1j * np.outer(tau, w))  # (i, w)
    M = meas.M[spin]
    for i in range(n):
        for j in range(n):
            out[meas.site[spin][i], meas.site[spin][j], meas.band[spin][i], meas.band[spin][j]] += \
                M[i, j] * np.outer(ph[i], ph[j].conj())
    return out


def mtilde_direct(geom: geo.Geometry, meas, spin: int) -> np.ndarray:
    """M~[b1, b2, k1, k2, w1E, w2E] = (1/Nc) sum_{ij} [b_i=b1][b_j=b2] exp(i(k1.r_i + w1 t_i))
    M_ij exp(-i(k2.r_j + w2 t_j)); band displacement phases are 1."""
    nb, nc, nE = geom.nb, geom.nc, geom.n_ext
    w = geo.vertex_frequencies(nE, geom.beta)
    n = meas.n[spin]
    out = np.zeros((nb, nb, nc, nc, nE, nE), complex)
    if n == 0:
        return out
    T = geom.cluster.dft  # (k, r)
    tau = meas.tau[spin]
    site, band = meas.site[spin], meas.band[spin]
    # Phi[b, k, w, i] = [b_i = b] exp(i k.r_i) exp(i w tau_i)
    phi = np.zeros((nb, nc, nE, n), complex)
    for i in range(n):
        phi[band[i], :, :, i] = np.outer(T[:, site[i]], np.exp(1j * w * tau[i]))
    phi2 = phi.reshape(nb * nc * nE, n)
    mt = phi2 @ meas.M[spin] @ phi2.conj().T / nc  # rows (b1,k1,w1) C-order, cols (b2,k2,w2)
    mt = mt.reshape(nb, nc, nE, nb, nc, nE)
    return np.transpose(mt, (0, 3, 1, 4, 2, 5))


def g_direct(geom: geo.Geometry, meas) -> np.ndarray:
    """G in the reference layout (b1, b2, s, k1, k2, w1E, w2E)."""
    nb, nc, nE = geom.nb, geom.nc, geom.n_ext
    g0 = geom.g0  # (b1, b2, s, k, w)
    G = np.zeros((nb, nb, 2, nc, nc, nE, nE), complex)
    for s in range(2):
        mt = mtilde_direct(geom, meas, s)  # (b1, b2, k1, k2, w1, w2)
        # -sum_{e,f} G0(b1,e,k1,w1) M~(e,f,k1,k2,w1,w2) G0(f,b2,k2,w2)
        G[:, :, s] = -np.einsum("aekw,efklwv,fblv->abklwv", g0[:, :, s], mt, g0[:, :, s])
        for k in range(nc):
            for wi in range(nE):
                G[:, :, s, k, k, wi, wi] += geom.beta * g0[:, :, s, k, wi]
    return G


def logical_G(geom: geo.Geometry, G: np.ndarray) -> np.ndarray:
    """L[s, b1, b2, k1, k2, w1E, w2E] = G(b2, b1, s, k1, k2, w1E, w2E): the value the reference's
    getGMultiband returns (tp_accumulator_cpu.hpp:388-405)."""
    return np.transpose(G, (2, 1, 0, 3, 4, 5, 6))


def g4_free(geom: geo.Geometry, channel: int, sign: float) -> np.ndarray:
    """G4 of one measurement with M = 0 on both spins, from closed forms, shaped
    (b1, b2, b3, b4, k1, w1, k2, w2, k_ex, w_ex) (Fortran order). With M = 0,
    G_s(b,b';k1,k2;w1,w2) = beta G0(b',b;k1,w1) delta_{k1k2} delta_{w1w2} (getGMultiband's
    band-transposed read), G0 spin independent, so (a = sign/2, sum over both spins):
      PP      :  sign beta^2 G0(b3,b1;k,w) G0(b4,b2;q-k,v-w)                    on k1=k2, w1=w2
      MAG     : -sign beta^2 G0(b1,b3;k,w) G0(b4,b2;k+q,w+v)                    on k1=k2, w1=w2
      CHARGE  : -sign beta^2 G0(b4,b1;k,w) G0(b3,b2;k+q,w+v)                    on k1=k2, w1=w2
                + 2 sign beta^2 G0(b1,b2;k1,w1) G0(b4,b3;k2,w2)                 if q = 0, v = 0
      TRANSV. : -sign beta^2 G0(b4,b2;k,w) G0(b1,b3;k+q,w+v)                    on k1=k2, w1=w2
    (TRANSVERSE from -a sum_s G_s(b2,b4;k1,k2) G_{-s}(b3,b1;k2+q,k1+q), tp_accumulator_cpu.hpp:454-489.)
    """
    nb, nc, nw = geom.nb, geom.nc, geom.n_freq
    ext = geom.ext
    add, sub = geom.cluster.add, geom.cluster.sub
    g0 = geom.g0[:, :, 0]  # spin independent (b, b', k, wE)
    b2 = geom.beta ** 2
    out = np.zeros(geom.g4_shape(), complex)
    gamma = geom.cluster.k_index((0.0, 0.0))
    ks = np.arange(nc)
    ws = np.arange(nw)
    g1 = g0[:, :, :, ext:ext + nw]  # (b, b', k, w) at the compact frequencies
    for iv, nu in enumerate(geom.w_ex):
        for iq, q in enumerate(geom.k_ex):
            if channel == geo.PARTICLE_PARTICLE_UP_DOWN:
                g2 = g0[:, :, sub[ks, q]][:, :, :, nu + nw - 1 - ws + ext]   # (b, b', k, w)
                val = sign * b2 * np.einsum("cakw,dbkw->abcdkw", g1, g2)
            elif channel == geo.PARTICLE_HOLE_MAGNETIC:
                g2 = g0[:, :, add[ks, q]][:, :, :, ws + nu + ext]
                val = -sign * b2 * np.einsum("ackw,dbkw->abcdkw", g1, g2)
            elif channel == geo.PARTICLE_HOLE_CHARGE:
                g2 = g0[:, :, add[ks, q]][:, :, :, ws + nu + ext]
                val = -sign * b2 * np.einsum("dakw,cbkw->abcdkw", g1, g2)
            elif channel == geo.PARTICLE_HOLE_TRANSVERSE:
                g2 = g0[:, :, add[ks, q]][:, :, :, ws + nu + ext]
                val = -sign * b2 * np.einsum("dbkw,ackw->abcdkw", g1, g2)
            else:
                raise ValueError("no closed form for this channel")
            for k in range(nc):
                out[:, :, :, :, k, :, k, :, iq, iv][..., ws, ws] += val[..., k, :]
            if channel == geo.PARTICLE_HOLE_CHARGE and q == gamma and nu == 0:
                out[..., iq, iv] += 2 * sign * b2 * np.einsum("abkw,dclv->abcdkwlv", g1, g1)
    return out


def difference(ref: np.ndarray, x: np.ndarray) -> dict:
    """l1 / l2 / l_inf relative differences (include/dca/function/util/difference.hpp:71-106)."""
    ref = np.asarray(ref).ravel()
    x = np.asarray(x).ravel()
    a = np.abs(ref)
    e = np.abs(ref - x)
    return {
        "l1": float(e.sum() / a.sum()) if a.sum() else float(e.sum()),
        "l2": float(np.sqrt((e ** 2).sum() / (a ** 2).sum())) if a.sum() else float(np.sqrt((e ** 2).sum())),
        "l_inf": float(e.max() / a.max()) if a.max() else float(e.max()),
    }
