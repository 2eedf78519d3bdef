"""Direct-definition evaluators for tiny inputs (TEST INFRASTRUCTURE ONLY).

These do not follow the reference's algorithm (no sorting, no per-orbital GEMMs, no scatter):
they evaluate the defining sums, so they pin the oracle's NDFT/DFT/G steps independently.

* NDFT definition used by the reference's own tests
  (test/unit/.../accumulation/single_sector_accumulation_test.hpp:226-254,
   test/unit/.../accumulation/tp/ndft/cached_ndft_test.hpp:120-143):
      f(b1, b2, r1, r2, w1, w2) = sum_{i in (b1,r1), j in (b2,r2)} M_ij exp(i (t_i w1 - t_j w2))
* Cluster DFT (space_transform_2D.hpp:41-44): f(k1, k2) = sum_{r1,r2} exp(i(k1 r1 - k2 r2)) f(r1,r2) / Nc
* G = -G0 M G0 + beta G0 delta (tp_accumulator_cpu.hpp:308-375) as a plain band-matrix product.
* Closed forms of G4 for M = 0 (G = beta G0 delta), with a spin-independent or Zeeman-split G0.
* G of a single vertex per spin (diagonal plus a non-symmetric rank-one part) in closed form, and
  the G4 channel definitions evaluated on a given G (g4_definition).
"""
from __future__ import annotations

import numpy as np

from dca_amd import geometry as geo


def column_labels(meas, spin: int):
    """(sites, bands) that label the columns of M: the vertices' left labels for non-density-density
    vertices (M is an inverse matrix, cached_ndft_base.hpp:97-103), else the same as the rows'."""
    if getattr(meas, "site_l", None) is not None:
        return meas.site_l[spin], meas.band_l[spin]
    return meas.site[spin], meas.band[spin]


def ndft_direct(geom: geo.Geometry, meas, spin: int) -> np.ndarray:
    """Returns f[r1, r2, b1, b2, w1E, w2E] by the defining double sum (rows of M at the vertices'
    right labels, columns at their left labels; M real or complex)."""
    nb, nc, nE = geom.nb, geom.nc, geom.n_ext
    w = geo.vertex_frequencies(nE, geom.beta)
    n = meas.n[spin]
    out = np.zeros((nc, nc, nb, nb, nE, nE), complex)
    if n == 0:
        return out
    tau = meas.tau[spin]
    ph = np.exp(1j * np.outer(tau, w))  # (i, w)
    M = meas.M[spin]
    site_c, band_c = column_labels(meas, spin)
    for i in range(n):
        for j in range(n):
            out[meas.site[spin][i], site_c[j], meas.band[spin][i], band_c[j]] += \
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

    def phi_of(site, band):
        # Phi[b, k, w, i] = [b_i = b] exp(i k.r_i) exp(i w tau_i)
        phi = np.zeros((nb, nc, nE, n), complex)
        for i in range(n):
            phi[band[i], :, :, i] = np.outer(T[:, site[i]], np.exp(1j * w * tau[i]))
        return phi.reshape(nb * nc * nE, n)

    rows = phi_of(meas.site[spin], meas.band[spin])
    cols = phi_of(*column_labels(meas, spin))
    mt = rows @ meas.M[spin] @ cols.conj().T / nc  # rows (b1,k1,w1) C-order, cols (b2,k2,w2)
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
    G_s(b,b';k1,k2;w1,w2) = beta g_s(b',b;k1,w1) delta_{k1k2} delta_{w1w2} (getGMultiband's
    band-transposed read of G_ = beta G0 delta), g_s = G0(., ., s, k, w), so with a = sign/2,
    sigma_0 = +1, sigma_1 = -1 and p + Q = (k + q, w + v), Q - p = (q - k, v - w):
      PP      :  a beta^2 sum_s g_s(b3,b1;p) g_-s(b4,b2;Q-p)                        on p1 = p2
      MAG     : -a beta^2 sum_s g_s(b1,b3;p) g_s(b4,b2;p+Q)                         on p1 = p2
                + a beta^2 [sum_s sigma_s g_s(b1,b2;p1)] [sum_s sigma_s g_s(b4,b3;p2)]  if Q = 0
      CHARGE  : -a beta^2 sum_s g_s(b4,b1;p) g_s(b3,b2;p+Q)                         on p1 = p2
                + a beta^2 [sum_s g_s(b1,b2;p1)] [sum_s g_s(b4,b3;p2)]                if Q = 0
      TRANSV. : -a beta^2 sum_s g_s(b4,b2;p) g_-s(b1,b3;p+Q)                        on p1 = p2
    from the channel definitions of tp_accumulator_cpu.hpp:454-661 (the pp, transverse and
    crossed terms pair one factor per spin; the magnetic direct term weighs the spins by
    sigma_s1 sigma_s2). A spin-independent G0 makes the magnetic direct term vanish and the s / -s
    pairings coincide; a Zeeman-split G0 (geometry.free_g0) keeps them apart.
    """
    nb, nc, nw = geom.nb, geom.nc, geom.n_freq
    ext = geom.ext
    add, sub = geom.cluster.add, geom.cluster.sub
    b2 = geom.beta ** 2
    a = 0.5 * sign
    out = np.zeros(geom.g4_shape(), complex)
    gamma = geom.cluster.k_index((0.0, 0.0))
    ks = np.arange(nc)
    ws = np.arange(nw)
    g = [geom.g0[:, :, s] for s in range(2)]            # (b, b', k, wE) per spin
    g1 = [x[:, :, :, ext:ext + nw] for x in g]          # (b, b', k, w) at the compact frequencies
    sigma = (1.0, -1.0)
    for iv, nu in enumerate(geom.w_ex):
        for iq, q in enumerate(geom.k_ex):
            val = np.zeros((nb, nb, nb, nb, nc, nw), complex)
            for s in range(2):
                t = 1 - s
                if channel == geo.PARTICLE_PARTICLE_UP_DOWN:
                    g2 = g[t][:, :, sub[ks, q]][:, :, :, nu + nw - 1 - ws + ext]
                    val += a * b2 * np.einsum("cakw,dbkw->abcdkw", g1[s], g2)
                elif channel == geo.PARTICLE_HOLE_MAGNETIC:
                    g2 = g[s][:, :, add[ks, q]][:, :, :, ws + nu + ext]
                    val -= a * b2 * np.einsum("ackw,dbkw->abcdkw", g1[s], g2)
                elif channel == geo.PARTICLE_HOLE_CHARGE:
                    g2 = g[s][:, :, add[ks, q]][:, :, :, ws + nu + ext]
                    val -= a * b2 * np.einsum("dakw,cbkw->abcdkw", g1[s], g2)
                elif channel == geo.PARTICLE_HOLE_TRANSVERSE:
                    g2 = g[t][:, :, add[ks, q]][:, :, :, ws + nu + ext]
                    val -= a * b2 * np.einsum("dbkw,ackw->abcdkw", g1[s], g2)
                else:
                    raise ValueError("no closed form for this channel")
            for k in range(nc):
                out[:, :, :, :, k, :, k, :, iq, iv][..., ws, ws] += val[..., k, :]
            if q == gamma and nu == 0 and channel in (geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE):
                wgt = sigma if channel == geo.PARTICLE_HOLE_MAGNETIC else (1.0, 1.0)
                d = wgt[0] * g1[0] + wgt[1] * g1[1]
                out[..., iq, iv] += a * b2 * np.einsum("abkw,dclv->abcdkwlv", d, d)
    return out


def g_single_vertex(geom: geo.Geometry, vertices) -> np.ndarray:
    """G in the reference layout (b1, b2, s, k1, k2, w1E, w2E) for an M with at most one vertex
    per spin sector, from the closed form of its NDFT: a vertex (tau, r, b) with M = (m) gives
    M~_s(b1,b2;k1,k2;w1,w2) = (m / Nc) [b1 = b][b2 = b] exp(i (k1 - k2).r + i (w1 - w2) tau), so
      G_s(b1, b2) = beta G0_s(b1,b2;p1) delta_{p1p2}
                    - (m / Nc) G0_s(b1,b;p1) exp(i(k1.r + w1 tau)) G0_s(b,b2;p2) exp(-i(k2.r + w2 tau)),
    a diagonal plus a rank-one matrix that is not symmetric, so every off-diagonal G read of the
    G4 update sees its argument order. vertices[s] = (tau, site, band, m) or None, or
    (tau, site, band, m, site_l, band_l) for a non-density-density vertex, whose column factor
    G0_s(b_l, b2; p2) exp(-i(k2.r_l + w2 tau)) carries its left labels; m may be complex."""
    nb, nc, nE = geom.nb, geom.nc, geom.n_ext
    w = geo.vertex_frequencies(nE, geom.beta)
    G = np.zeros((nb, nb, 2, nc, nc, nE, nE), complex)
    for s in range(2):
        g0 = geom.g0[:, :, s]  # (b, b', k, w)
        for k in range(nc):
            for i in range(nE):
                G[:, :, s, k, k, i, i] += geom.beta * g0[:, :, k, i]
        if vertices[s] is None:
            continue
        tau, site, band, m = vertices[s][:4]
        site_l, band_l = vertices[s][4:6] if len(vertices[s]) > 4 else (site, band)

        def phase(r):
            return np.exp(1j * (geom.cluster.k @ geom.cluster.r[r])[:, None] + 1j * w[None, :] * tau)  # (k, w)

        left = g0[:, band] * phase(site)[None]                 # (b1, k1, w1)
        right = g0[band_l, :] * np.conj(phase(site_l))[None]   # (b2, k2, w2)
        G[:, :, s] -= (m / nc) * np.einsum("akw,blv->abklwv", left, right)
    return G


def g4_definition(geom: geo.Geometry, G: np.ndarray, channel: int, sign: complex,
                  variant: str = "", spin_symmetric: bool = True) -> np.ndarray:
    """G4 of one measurement from its G (reference layout), by the channel definitions, shaped
    like g4_free. With L_s(b, b'; p1, p2) = G(b', b, s, k1, k2, w1 + ext, w2 + ext) (getGMultiband's
    band-transposed read, tp_accumulator_cpu.hpp:388-405), a = sign / 2, p + Q = (k + q, w + v) and
    Q - p = (q - k, v + nw - 1 - w):
      MAG    : a sum_{s1,s2} sigma_s1 sigma_s2 L_s1(b2,b1; p1, p1+Q) L_s2(b3,b4; p2+Q, p2)
               - a sum_s L_s(b3,b1; p1, p2) L_s(b2,b4; p2+Q, p1+Q)
      CHARGE : a sum_{s1,s2} L_s1(b2,b1; p1, p1+Q) L_s2(b3,b4; p2+Q, p2)
               - a sum_s L_s(b1,b4; p1, p2) L_s(b2,b3; p2+Q, p1+Q)
      TRANSV.: -a sum_s L_s(b2,b4; p1, p2) L_-s(b3,b1; p2+Q, p1+Q)
      PP     : a sum_s L_s(b1,b3; p1, p2) L_-s(b2,b4; Q-p1, Q-p2)
    For one band the momentum and frequency arguments of the ph channels are the Wick
    contractions written in the reference's comments (tp_accumulator_cpu.hpp:455-456, 491-495,
    549-553): <c+(k1+q) c(k1) c+(k2) c(k2+q)> = G(k1, k1+q) G(k2+q, k2) - G(k2+q, k1+q) G(k1, k2)
    with G(x, y) ~ <c(x) c+(y)>. For pp the comment (:629-630) reads G(q-k2, q-k1, s) G(k2, k1, -s),
    the transpose of both factors; the code (:640-648), and so every result of the reference,
    contracts G(k1, k2, s) G(q-k1, q-k2, -s), which is what is written here. The band slots of the
    several-band terms are those of the code (:513-523, 532-543, 561-572, 640-648).

    Models that are not spin symmetric (spin_symmetric=False; the spin is part of the band index)
    read spin sector 0 only, with the whole sign and no 1/2 (tp_accumulator_cpu.hpp:668-757):
      PP     : sign [L_0(b1,b3; p1, p2) L_0(b2,b4; Q-p1, Q-p2) - L_0(b1,b4; p1, Q-p2) L_0(b2,b3; Q-p1, p2)]
      MAG    : sign [L_0(b2,b1; p1, p1+Q) L_0(b3,b4; p2+Q, p2) - L_0(b3,b1; p2+Q, p1+Q) L_0(b2,b4; p1, p2)]
    the Wick contractions of <c+(k1+q, s1) c(k1, s2) c+(k2, s4) c(k2+q, s3)> written at :711-714 and
    their pp counterpart (:681-683, :694-695; as in the symmetric pp term, the code's argument
    order). The other channels are not defined there (:752-753).

    variant (test-only planted mistakes, to show that the checks see them): "pp_comment" (the pp
    comment's argument order), "crossed_transposed" (G(p1+Q, p2+Q) in the crossed ph term),
    "transverse_same_spin" (s instead of -s), "mag_spin_sum" (no sigma weights in the magnetic
    direct term), "nss_pp_exchange_slots" (b3 <-> b4 in the exchange term of the non-symmetric pp),
    "nss_half" (the symmetric branch's 1/2 in the non-symmetric branch)."""
    nb, nc, nw = geom.nb, geom.nc, geom.n_freq
    ext = geom.ext
    add, sub = geom.cluster.add, geom.cluster.sub
    L = np.transpose(G, (2, 1, 0, 3, 4, 5, 6))  # (s, b, b', k1, k2, w1E, w2E)
    a = 0.5 * sign
    sigma = (1.0, -1.0) if variant != "mag_spin_sum" else (1.0, 1.0)
    K1 = np.arange(nc)[:, None, None, None]
    W1 = np.arange(nw)[None, :, None, None]
    K2 = np.arange(nc)[None, None, :, None]
    W2 = np.arange(nw)[None, None, None, :]

    def read(s, kA, kB, wA, wB):
        kA, kB, wA, wB = np.broadcast_arrays(kA, kB, wA, wB)
        return L[s][:, :, kA, kB, wA + ext, wB + ext]  # (b, b', k1, w1, k2, w2)

    out = np.zeros(geom.g4_shape(), complex)
    if not spin_symmetric:
        f = sign * (0.5 if variant == "nss_half" else 1.0)
        for iv, nu in enumerate(geom.w_ex):
            for iq, q in enumerate(geom.k_ex):
                K1q, K2q = add[K1, q], add[K2, q]
                mK1, mK2 = sub[K1, q], sub[K2, q]
                mW1, mW2 = nu + nw - 1 - W1, nu + nw - 1 - W2
                if channel == geo.PARTICLE_PARTICLE_UP_DOWN:
                    val = np.einsum("ac...,bd...->abcd...", read(0, K1, K2, W1, W2), read(0, mK1, mK2, mW1, mW2))
                    ex = "ac...,bd...->abcd..." if variant == "nss_pp_exchange_slots" else "ad...,bc...->abcd..."
                    val = val - np.einsum(ex, read(0, K1, mK2, W1, mW2), read(0, mK1, K2, mW1, W2))
                elif channel == geo.PARTICLE_HOLE_MAGNETIC:
                    val = np.einsum("ba...,cd...->abcd...", read(0, K1, K1q, W1, W1 + nu), read(0, K2q, K2, W2 + nu, W2))
                    val = val - np.einsum("ca...,bd...->abcd...", read(0, K2q, K1q, W2 + nu, W1 + nu),
                                          read(0, K1, K2, W1, W2))
                else:
                    raise ValueError("channel not defined for models that are not spin symmetric")
                out[..., iq, iv] = f * val
        return out
    for iv, nu in enumerate(geom.w_ex):
        for iq, q in enumerate(geom.k_ex):
            K1q, K2q = add[K1, q], add[K2, q]
            val = np.zeros((nb, nb, nb, nb, nc, nw, nc, nw), complex)
            if channel in (geo.PARTICLE_HOLE_MAGNETIC, geo.PARTICLE_HOLE_CHARGE):
                wgt = sigma if channel == geo.PARTICLE_HOLE_MAGNETIC else (1.0, 1.0)
                A = sum(wgt[s] * read(s, K1, K1q, W1, W1 + nu) for s in range(2))
                B = sum(wgt[s] * read(s, K2q, K2, W2 + nu, W2) for s in range(2))
                val += a * np.einsum("ba...,cd...->abcd...", A, B)
                for s in range(2):
                    X = read(s, K1, K2, W1, W2)
                    if variant == "crossed_transposed":
                        Y = read(s, K1q, K2q, W1 + nu, W2 + nu)
                    else:
                        Y = read(s, K2q, K1q, W2 + nu, W1 + nu)
                    if channel == geo.PARTICLE_HOLE_MAGNETIC:
                        val -= a * np.einsum("ca...,bd...->abcd...", X, Y)
                    else:
                        val -= a * np.einsum("ad...,bc...->abcd...", X, Y)
            elif channel == geo.PARTICLE_HOLE_TRANSVERSE:
                for s in range(2):
                    t = s if variant == "transverse_same_spin" else 1 - s
                    X = read(s, K1, K2, W1, W2)
                    Y = read(t, K2q, K1q, W2 + nu, W1 + nu)
                    val -= a * np.einsum("bd...,ca...->abcd...", X, Y)
            elif channel == geo.PARTICLE_PARTICLE_UP_DOWN:
                mK1, mK2 = sub[K1, q], sub[K2, q]
                mW1, mW2 = nu + nw - 1 - W1, nu + nw - 1 - W2
                for s in range(2):
                    if variant == "pp_comment":
                        X = read(s, mK2, mK1, mW2, mW1)
                        Y = read(1 - s, K2, K1, W2, W1)
                    else:
                        X = read(s, K1, K2, W1, W2)
                        Y = read(1 - s, mK1, mK2, mW1, mW2)
                    val += a * np.einsum("ac...,bd...->abcd...", X, Y)
            else:
                raise ValueError("unknown channel")
            out[..., iq, iv] = val
    return out


def compute_m(N: np.ndarray, exp_v: np.ndarray) -> np.ndarray:
    """The M a CT-AUX walker hands to the accumulator: M = diag(exp(V) - 1) N, row i scaled by
    exp(V_i) - 1 in FP64 (CtauxWalker::computeM, ctaux_walker.hpp:1889-1916, with
    multiplyDiagonalLeft, src/linalg/lapack/multiply_diagonal_gpu.cu:33-41)."""
    N = np.asarray(N, np.float64)
    return np.asfortranarray((np.asarray(exp_v, np.float64) - 1.0)[:, None] * N)


def normalize_g4(g4: np.ndarray, beta: float, accumulated_sign: float) -> np.ndarray:
    """The solvers' G4 normalisation after the accumulation: G4 / (accumulated_sign beta^2)
    (ctaux_cluster_solver.hpp:523-527, ctint_cluster_solver.hpp:511-516), in FP64."""
    return np.asarray(g4, complex) / (accumulated_sign * beta * beta)


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
