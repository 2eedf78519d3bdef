"""CPU restatement of the reference's single-particle accumulator (TEST INFRASTRUCTURE ONLY).

SpAccumulator<Parameters, CPU> (include/dca/phys/dca_step/cluster_solver/shared_tools/
accumulation/sp/sp_accumulator.hpp:160-237) with its delayed NFFT
Dnfft1D<double, WDmn, PDmn, oversampling = 8, CUBIC> (include/dca/math/nfft/dnfft_1d.hpp), the
Kaiser-Bessel window (src/math/nfft/window_functions/kaiser_bessel_function.cpp) and the time
domains (include/dca/math/nfft/domains/nfft_time_domain.hpp:118-200), in numpy. The final FFT
that the reference does with FFTW's r2c plan (dnfft_1d.hpp:437-500) is numpy.fft.rfft, which
computes the same forward transform sum_t f[t] exp(-2 pi i k t / n).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.

What the delayed NFFT approximates (and what pins this restatement, tests/test_sp.py): for real M,
    M_r_w(b1, s, b2, s, r, w) = (sign / Nc) sum_{i, j} M_s(i, j) exp(i w (tau_i - tau_j))
                                [b_i = b1, b_j = b2, r_i - r_j = r]
over the single-particle Matsubara frequencies w = (2l + 1) pi / beta
(src/phys/domains/time_and_frequency/frequency_domain.cpp:29-49).
"""
from __future__ import annotations

import math

import numpy as np

OVERSAMPLING = 8        # SpAccumulator::oversampling (sp_accumulator.hpp:62)
WINDOW_SAMPLING = 32    # Dnfft1D::window_sampling_ (dnfft_1d.hpp:88)
WINDOW_SIGMA = 2.0      # Dnfft1D::window_function_sigma_ (dnfft_1d.hpp:89)


def sp_frequency_indices(n_sp_freq: int) -> np.ndarray:
    """frequency_domain indices (frequency_domain.cpp:29-49): +-(2l + 1), negative half first."""
    size = 2 * n_sp_freq
    idx = np.empty(size, np.int64)
    for l in range(n_sp_freq):
        idx[size // 2 + l] = 2 * l + 1
        idx[size // 2 - 1 - l] = -(2 * l + 1)
    return idx


class KaiserBessel:
    """kaiser_bessel_function (kaiser_bessel_function.cpp) with n, m, sigma as Dnfft1D sets them
    (dnfft_1d.hpp:202-205)."""

    def __init__(self, n: int, m: int, sigma: float):
        self.n, self.m, self.sigma = n, m, sigma

    def b_val(self):
        return math.pi * (2.0 - 1.0 / self.sigma)

    def renorm(self):
        return 1.0 / (math.pi * self.m) * math.sinh(math.pi * self.m * (2.0 * self.sigma - 1.0) / self.sigma)

    def phi_t(self, x: float) -> float:
        val = self.m * self.m - self.n * self.n * x * x
        b = self.b_val()
        if val > 1e-12:
            r = 1.0 / math.pi * math.sinh(b * math.sqrt(val)) / math.sqrt(val)
        elif val < -1e-12:
            r = 1.0 / math.pi * math.sin(b * math.sqrt(-val)) / math.sqrt(-val)
        else:
            r = b / math.pi
        return r / self.renorm()

    def d_phi_t(self, x: float) -> float:
        val = self.m * self.m - self.n * self.n * x * x
        b, n = self.b_val(), self.n
        if val > 1e-12:
            r = 1.0 / math.pi * (n * n * x) * ((-b * math.cosh(b * math.sqrt(val))) / val +
                                               math.sinh(b * math.sqrt(val)) / val ** 1.5)
        elif val < -1e-12:
            r = 1.0 / math.pi * (n * n * x) * ((b * math.cos(b * math.sqrt(-val))) / (-val) -
                                               math.sin(b * math.sqrt(-val)) / (-val) ** 1.5)
        else:
            return 0.0
        return r / self.renorm()

    @staticmethod
    def bessel_i0(x: float) -> float:
        """Power series of kaiser_bessel_function::besselI0 (terms down to 1e-16)."""
        result, sq, gamma, y, l = 0.0, x * x / 4.0, 1.0, 1.0, 0
        while True:
            term = y / gamma
            if abs(term) < 1e-16:
                break
            result += term
            gamma *= (l + 1) * (l + 1)
            y *= sq
            l += 1
        return result

    def phi_wn(self, wn: int) -> float:
        n, m = self.n, self.m
        if wn * wn >= (n * (1.0 - 1.0 / (2.0 * self.sigma))) ** 2:
            raise ValueError("phi_wn outside the window's band")
        arg = m * math.sqrt(self.b_val() ** 2 - (2.0 * math.pi * wn / n) ** 2)
        return 1.0 / n * self.bessel_i0(arg) / self.renorm()


class Dnfft1D:
    """Dnfft1D<double, WDmn, PDmn, oversampling, CUBIC> for real samples (dnfft_1d.hpp)."""

    def __init__(self, n_sp_freq: int, n_labels: int, oversampling: int = OVERSAMPLING):
        os_, ws = oversampling, WINDOW_SAMPLING
        self.os, self.ws = os_, ws
        mf = n_sp_freq                                   # maximumFrequency() = WDmn size / 2
        self.n = os_ * mf                                # LEFT_ORIENTED size
        self.n_pad = self.n + 2 * os_                    # PADDED size
        self.Delta = 1.0 / (os_ * mf)
        self.delta = 1.0 / (os_ * mf * ws)
        self.one_div_Delta = 1.0 / self.Delta
        self.one_div_delta = 1.0 / self.delta
        self.T0 = -0.5 - os_ * self.Delta                # PADDED first element
        self.t0 = -2.0 / mf                              # WINDOW_FUNCTION first element
        # Window function over the padded time domain (dnfft_1d.hpp:202-205).
        self.kb = KaiserBessel(self.n_pad, os_, WINDOW_SIGMA)
        # Cubic coefficients on the fine grid (dnfft_1d.hpp:207-250): convolution_times(i, j) is
        # WINDOW_FUNCTION element j + i ws (i < 4 os, j < ws).
        fine = self.t0 + np.arange(4 * os_ * ws) * self.delta
        conv = fine.reshape(4 * os_, ws)
        d = conv[0, 1] - conv[0, 0]
        cub = np.zeros((4, 4 * os_, ws))
        for i in range(4 * os_):
            for j in range(ws):
                tau = conv[i, j]
                f0, f1 = self.kb.phi_t(tau), self.kb.phi_t(tau + d)
                df0, df1 = self.kb.d_phi_t(tau), self.kb.d_phi_t(tau + d)
                cub[0, i, j] = f0
                cub[1, i, j] = df0
                cub[2, i, j] = -(3.0 * f0 - 3.0 * f1 + 2.0 * df0 * d + df1 * d) / d ** 2
                cub[3, i, j] = -(-2.0 * f0 + 2.0 * f1 - 1.0 * df0 * d - df1 * d) / d ** 3
        self.cubic = cub
        self.freq_idx = sp_frequency_indices(n_sp_freq)
        self.phi_wn = np.array([self.kb.phi_wn(int(k)) for k in self.freq_idx])
        self.f_tau = np.zeros((self.n_pad, n_labels))

    def accumulate(self, labels: np.ndarray, t_val: np.ndarray, f_val: np.ndarray):
        """convoluteToFTauFineCubicInterpolation (dnfft_1d.hpp:358-399) for every sample, in the
        given order."""
        os_, ws = self.os, self.ws
        for lab, t, f in zip(labels, t_val, f_val):
            tau_0 = int((t - self.T0) * self.one_div_Delta)
            t0_lb = self.T0 + tau_0 * self.Delta
            tau_1 = int((t0_lb - t - self.t0) * self.one_div_delta)
            t1_lb = self.t0 + tau_1 * self.delta
            diff = t0_lb - t - t1_lb
            y = np.array([f, f * diff, f * diff * diff, f * diff * diff * diff])
            tau_index = tau_0 - os_
            dti = tau_1 - os_ * ws
            j = dti % ws
            i = (dti - j) // ws
            # NfftAtomicConvolution<os>::execute_cubic: f[count] for count = 1 .. 2 os
            for count in range(1, 2 * os_ + 1):
                self.f_tau[tau_index + count, lab] += (self.cubic[0, i + count, j] * y[0] +
                                                       self.cubic[1, i + count, j] * y[1] +
                                                       self.cubic[2, i + count, j] * y[2] +
                                                       self.cubic[3, i + count, j] * y[3])

    def finalize(self) -> np.ndarray:
        """foldTimeDomainBack + transformFTauToFW for real samples (dnfft_1d.hpp:401-500);
        returns f_w[w, label]."""
        n, pad = self.n, self.os
        f = self.f_tau.copy()
        f[n:n + pad] += f[:pad]                          # fold the left halo
        f[self.n_pad - pad - n:self.n_pad - n] += f[self.n_pad - pad:]   # fold the right halo
        f_in = f[pad:pad + n]
        out = np.fft.rfft(f_in, axis=0)                  # (n/2 + 1, labels)
        omega = np.empty((n, f_in.shape[1]), complex)
        t = np.arange(n // 2)
        omega[:n // 2] = -out[t].real + 1j * out[t].imag
        t = np.arange(n // 2, n)
        omega[n // 2:] = -out[n - t].real - 1j * out[n - t].imag
        w_idx = np.where(self.freq_idx >= 0, self.freq_idx, self.freq_idx + n)
        return omega[w_idx] / self.phi_wn[:, None] / n


class SpOracle:
    """SpAccumulator<CPU> restated: accumulate(M_pair, configs, factor), finalize,
    get_sign_times_M_r_w(). rsub[a, b] = index(r_b - r_a) is the cluster's site subtraction
    (RDmn::parameter_type::subtract)."""

    def __init__(self, nb: int, nc: int, beta: float, n_sp_freq: int, rsub: np.ndarray,
                 accumulate_m_sqr: bool = False):
        self.nb, self.nc, self.beta, self.n_sp_freq = nb, nc, beta, n_sp_freq
        self.rsub = np.asarray(rsub)
        self.sqr = accumulate_m_sqr
        self.resetAccumulation()

    def resetAccumulation(self):
        n_lab = self.nb * self.nb * self.nc
        self.nfft = [Dnfft1D(self.n_sp_freq, n_lab) for _ in range(2)]
        self.nfft_sqr = [Dnfft1D(self.n_sp_freq, n_lab) for _ in range(2)] if self.sqr else None

    def accumulate(self, meas, factor: float):
        """sp_accumulator.hpp:160-203: for each spin, j outer, i inner (density-density vertices:
        left and right orbitals agree)."""
        one_div_two_beta = 1.0 / (2.0 * self.beta)
        for s in range(2):
            n = meas.n[s]
            if n == 0:
                continue
            tau, site, band = meas.tau[s], meas.site[s], meas.band[s]
            M = meas.M[s]
            jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")  # j outer, i inner
            jj, ii = jj.ravel(), ii.ravel()
            delta_r = self.rsub[site[jj], site[ii]]
            lab = band[ii] + self.nb * (band[jj] + self.nb * delta_r)
            t = (tau[ii] - tau[jj]) * one_div_two_beta
            f = M[ii, jj]
            self.nfft[s].accumulate(lab, t, factor * f)
            if self.sqr:
                self.nfft_sqr[s].accumulate(lab, t, factor * f * f)

    def _finalize(self, objs) -> np.ndarray:
        """finalizeFunction (sp_accumulator.hpp:205-221): M_r_w(b1, s, b2, s, r, w) +=
        tmp(w, b1, b2, r) / Nc; returns the (nb, 2, nb, 2, nc, nW) array."""
        nb, nc = self.nb, self.nc
        nw = 2 * self.n_sp_freq
        out = np.zeros((nb, 2, nb, 2, nc, nw), complex)
        for s in range(2):
            fw = objs[s].finalize()                        # (w, label), label = b1 + nb (b2 + nb r)
            t = fw.reshape(nw, nc, nb, nb)                 # (w, r, b2, b1)
            out[:, s, :, s] += np.transpose(t, (3, 2, 1, 0)) / nc
        return out

    def get_sign_times_M_r_w(self) -> np.ndarray:
        return self._finalize(self.nfft)

    def get_sign_times_M_r_w_sqr(self) -> np.ndarray:
        return self._finalize(self.nfft_sqr)


def m_r_w_exact(nb: int, nc: int, beta: float, n_sp_freq: int, rsub: np.ndarray, measurements,
                squared: bool = False) -> np.ndarray:
    """The exact transform the delayed NFFT approximates:
    (1/Nc) sum_meas sign sum_{ij} M_ij exp(i w (tau_i - tau_j)) [b_i, b_j, r_i - r_j], shaped
    (b1, s, b2, s, r, w) like SpOracle's output."""
    w = sp_frequency_indices(n_sp_freq) * math.pi / beta
    out = np.zeros((nb, 2, nb, 2, nc, len(w)), complex)
    for meas in measurements:
        for s in range(2):
            n = meas.n[s]
            if n == 0:
                continue
            tau, site, band = meas.tau[s], meas.site[s], meas.band[s]
            M = meas.M[s] ** 2 if squared else meas.M[s]
            ph = np.exp(1j * np.outer(tau, w))                 # (i, w)
            for i in range(n):
                dr = rsub[site, site[i]]                        # index(r_i - r_j) for all j
                contrib = meas.sign * M[i, :, None] * ph[i][None, :] * np.conj(ph)   # (j, w)
                np.add.at(out, (band[i], s, band, s, dr), contrib)
    return out / nc
