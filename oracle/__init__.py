"""ctypes binding of the FP64 CPU oracle (oracle/tp_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg
of bench.py, as the checker. The product (dca_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libtp_oracle.so")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


class _Geometry(ctypes.Structure):
    _fields_ = [
        ("nb", ctypes.c_int), ("nc", ctypes.c_int), ("beta", ctypes.c_double),
        ("nw", ctypes.c_int),
        ("n_wex", ctypes.c_int), ("wex", ctypes.POINTER(ctypes.c_int)),
        ("n_kex", ctypes.c_int), ("kex", ctypes.POINTER(ctypes.c_int)),
        ("n_channels", ctypes.c_int), ("channels", ctypes.POINTER(ctypes.c_int)),
        ("k_add", ctypes.POINTER(ctypes.c_int)), ("k_sub", ctypes.POINTER(ctypes.c_int)),
        ("dft", ctypes.POINTER(ctypes.c_double)), ("band_phase", ctypes.POINTER(ctypes.c_double)),
        ("g0", ctypes.POINTER(ctypes.c_double)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.tpo_create.restype = vp
        L.tpo_create.argtypes = [ctypes.POINTER(_Geometry), ctypes.c_char_p, ctypes.c_int]
        L.tpo_destroy.argtypes = [vp]
        L.tpo_reset.argtypes = [vp]
        arr_i = ctypes.POINTER(ctypes.c_int)
        arr_pd = ctypes.POINTER(ctypes.c_void_p)
        L.tpo_accumulate.restype = ctypes.c_double
        L.tpo_accumulate.argtypes = [vp, arr_i, arr_pd, arr_i, arr_pd, arr_pd, arr_pd, ctypes.c_double]
        L.tpo_compute_G.argtypes = [vp, arr_i, arr_pd, arr_i, arr_pd, arr_pd, arr_pd]
        L.tpo_set_spin_symmetric.argtypes = [vp, ctypes.c_int]
        L.tpo_accumulate_general.restype = ctypes.c_double
        L.tpo_accumulate_general.argtypes = [vp, arr_i, arr_pd, ctypes.c_int, arr_i, arr_pd, arr_pd, arr_pd,
                                             arr_pd, arr_pd, ctypes.c_double, ctypes.c_double]
        L.tpo_compute_G_general.argtypes = [vp, arr_i, arr_pd, ctypes.c_int, arr_i, arr_pd, arr_pd, arr_pd,
                                            arr_pd, arr_pd]
        for f in (L.tpo_get_Mrrww, L.tpo_get_G):
            f.restype = ctypes.c_longlong
            f.argtypes = [vp, ctypes.c_void_p, ctypes.c_longlong]
        L.tpo_get_G4.restype = ctypes.c_longlong
        L.tpo_get_G4.argtypes = [vp, ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong]
        L.tpo_g4_size.restype = ctypes.c_longlong
        L.tpo_g4_size.argtypes = [vp]
        L.tpo_g_size.restype = ctypes.c_longlong
        L.tpo_g_size.argtypes = [vp]
        L.tpo_sum_to.restype = ctypes.c_int
        L.tpo_sum_to.argtypes = [vp, vp]
        L.tpo_keep_intermediate.argtypes = [vp, ctypes.c_int]
        L.tpo_use_blas.restype = ctypes.c_int
        L.tpo_use_blas.argtypes = [ctypes.c_char_p]
        L.tpo_blas_name.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _is_general(meas) -> bool:
    """A measurement of the general model: complex M or sign, or vertices with left labels."""
    return (getattr(meas, "site_l", None) is not None or any(np.iscomplexobj(x) for x in meas.M)
            or isinstance(meas.sign, complex))


BLAS_CANDIDATES = ("/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so", "libopenblas.so")


def use_blas(enable: bool = True) -> str:
    """Switches the oracle's complex GEMMs to BLAS zgemm_ (the first of BLAS_CANDIDATES that
    loads, single-threaded) or back to its own loop; returns the name of the GEMM in use."""
    L = lib()
    if not enable:
        L.tpo_use_blas(None)
    else:
        for path in BLAS_CANDIDATES:
            if L.tpo_use_blas(path.encode()):
                break
    return L.tpo_blas_name().decode()


class OracleAccumulator:
    """FP64 CPU TpAccumulator restatement (same method names as the reference class)."""

    def __init__(self, geom, spin_symmetric: bool = True):
        self.geom = geom
        self._keep = [
            np.ascontiguousarray(geom.w_ex, np.int32),
            np.ascontiguousarray(geom.k_ex, np.int32),
            np.ascontiguousarray(geom.channels, np.int32),
            np.ascontiguousarray(geom.cluster.add.ravel(order="F"), np.int32),
            np.ascontiguousarray(geom.cluster.sub.ravel(order="F"), np.int32),
            np.ascontiguousarray(geom.cluster.dft.ravel(order="F"), np.complex128),
            None if geom.band_phase is None else
            np.ascontiguousarray(geom.band_phase.ravel(order="F"), np.complex128),
            np.ascontiguousarray(geom.g0.ravel(order="F"), np.complex128),
        ]
        wex, kex, ch, add, sub, dft, bp, g0 = self._keep
        g = _Geometry()
        g.nb, g.nc, g.beta, g.nw = geom.nb, geom.nc, geom.beta, geom.n_freq
        g.n_wex, g.wex = len(wex), _ptr(wex, ctypes.c_int)
        g.n_kex, g.kex = len(kex), _ptr(kex, ctypes.c_int)
        g.n_channels, g.channels = len(ch), _ptr(ch, ctypes.c_int)
        g.k_add, g.k_sub = _ptr(add, ctypes.c_int), _ptr(sub, ctypes.c_int)
        g.dft = dft.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        g.band_phase = None if bp is None else bp.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        g.g0 = g0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        err = ctypes.create_string_buffer(256)
        self._h = lib().tpo_create(ctypes.byref(g), err, 256)
        if not self._h:
            raise ValueError(err.value.decode())
        lib().tpo_set_spin_symmetric(self._h, 1 if spin_symmetric else 0)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().tpo_destroy(self._h)
            self._h = None

    def resetAccumulation(self):
        lib().tpo_reset(self._h)

    def keep_intermediate(self, keep: bool):
        """False: the M_r_r_w_w copy for get_Mrrww is not kept (timing runs)."""
        lib().tpo_keep_intermediate(self._h, 1 if keep else 0)

    def _args(self, meas):
        keep = []
        n = (ctypes.c_int * 2)(*meas.n)
        ld = (ctypes.c_int * 2)(*[max(1, x) for x in meas.n])
        m = (ctypes.c_void_p * 2)()
        tau = (ctypes.c_void_p * 2)()
        site = (ctypes.c_void_p * 2)()
        band = (ctypes.c_void_p * 2)()
        for s in range(2):
            a = np.asfortranarray(meas.M[s], np.float64)
            t = np.ascontiguousarray(meas.tau[s], np.float64)
            r = np.ascontiguousarray(meas.site[s], np.int32)
            b = np.ascontiguousarray(meas.band[s], np.int32)
            keep += [a, t, r, b]
            m[s], tau[s], site[s], band[s] = a.ctypes.data, t.ctypes.data, r.ctypes.data, b.ctypes.data
        return keep, (n, m, ld, tau, site, band)

    def _general_args(self, meas):
        """Arguments of the general entry points: complex M when any sector is complex, the left
        labels (site_l, band_l) when the measurement carries them."""
        keep = []
        cplx = any(np.iscomplexobj(x) for x in meas.M)
        n = (ctypes.c_int * 2)(*meas.n)
        ld = (ctypes.c_int * 2)(*[max(1, x) for x in meas.n])
        m = (ctypes.c_void_p * 2)()
        tau = (ctypes.c_void_p * 2)()
        site = (ctypes.c_void_p * 2)()
        band = (ctypes.c_void_p * 2)()
        nd = getattr(meas, "site_l", None) is not None
        site_l = (ctypes.c_void_p * 2)() if nd else None
        band_l = (ctypes.c_void_p * 2)() if nd else None
        for s in range(2):
            a = np.asfortranarray(meas.M[s], np.complex128 if cplx else np.float64)
            t = np.ascontiguousarray(meas.tau[s], np.float64)
            r = np.ascontiguousarray(meas.site[s], np.int32)
            b = np.ascontiguousarray(meas.band[s], np.int32)
            keep += [a, t, r, b]
            m[s], tau[s], site[s], band[s] = a.ctypes.data, t.ctypes.data, r.ctypes.data, b.ctypes.data
            if nd:
                rl = np.ascontiguousarray(meas.site_l[s], np.int32)
                bl = np.ascontiguousarray(meas.band_l[s], np.int32)
                keep += [rl, bl]
                site_l[s], band_l[s] = rl.ctypes.data, bl.ctypes.data
        return keep, (n, m, int(cplx), ld, tau, site, band, site_l, band_l)

    def accumulate(self, meas) -> float:
        if _is_general(meas):
            keep, args = self._general_args(meas)
            sign = complex(meas.sign)
            f = lib().tpo_accumulate_general(self._h, *args, sign.real, sign.imag)
        else:
            keep, (n, m, ld, tau, site, band) = self._args(meas)
            f = lib().tpo_accumulate(self._h, n, m, ld, tau, site, band, float(meas.sign))
        if f < 0:
            raise NotImplementedError("Specified four point type not implemented.")
        return f

    def compute_G(self, meas):
        if _is_general(meas):
            keep, args = self._general_args(meas)
            lib().tpo_compute_G_general(self._h, *args)
            return
        keep, (n, m, ld, tau, site, band) = self._args(meas)
        lib().tpo_compute_G(self._h, n, m, ld, tau, site, band)

    def get_G(self) -> np.ndarray:
        """G in the reference layout (b1, b2, s, k1, k2, w1E, w2E), returned as an array of that
        shape (Fortran order)."""
        g = self.geom
        n = lib().tpo_g_size(self._h)
        out = np.empty(n, np.complex128)
        lib().tpo_get_G(self._h, out.ctypes.data, n)
        return out.reshape((g.nb, g.nb, 2, g.nc, g.nc, g.n_ext, g.n_ext), order="F")

    def get_Mrrww(self) -> np.ndarray:
        g = self.geom
        n = g.nc * g.nc * g.nb * g.nb * 2 * g.n_ext * g.n_ext
        out = np.empty(n, np.complex128)
        lib().tpo_get_Mrrww(self._h, out.ctypes.data, n)
        return out.reshape((g.nc, g.nc, g.nb, g.nb, 2, g.n_ext, g.n_ext), order="F")

    def get_G4(self) -> list[np.ndarray]:
        """One flat complex128 vector per channel, in the reference's G4 order
        (b1, b2, b3, b4, k1, w1, k2, w2, k_ex, w_ex), first index fastest."""
        n = lib().tpo_g4_size(self._h)
        res = []
        for c in range(len(self.geom.channels)):
            out = np.empty(n, np.complex128)
            if lib().tpo_get_G4(self._h, c, out.ctypes.data, n) < 0:
                raise RuntimeError("There is no G4 stored in this class.")
            res.append(out)
        return res

    def sumTo(self, other: "OracleAccumulator"):
        if lib().tpo_sum_to(self._h, other._h) != 0:
            raise ValueError("Objects accumulate different number of channels.")
