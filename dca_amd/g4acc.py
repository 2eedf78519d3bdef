"""Python mirror of DCA++'s TpAccumulator over the g4acc C ABI (include/g4acc.h).

Method names, argument meaning and error behaviour follow the reference class
(include/dca/phys/dca_step/cluster_solver/shared_tools/accumulation/tp/tp_accumulator_cpu.hpp:92-146,
 .../tp/tp_accumulator_gpu.hpp:103-187): resetAccumulation, accumulate(M_pair, configs, sign),
finalize, sumTo, get_G4, synchronizeCopy. Status codes become the reference's exception types:
std::logic_error -> LogicError, std::bad_alloc -> MemoryError, launch failures -> RuntimeError.

The compute path is the HIP library dca_amd/libg4acc.so; there is no CPU fallback: if the
library is missing, importing this module's loader raises.
"""
from __future__ import annotations

import ctypes
import os
import sys
from typing import Sequence

import numpy as np

from . import geometry as geo

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libg4acc.so")

OK, INVALID_ARG, UNSUPPORTED_CHANNEL, NOT_INITIALIZED, OUT_OF_MEMORY, DEVICE_ERROR, EMPTY = range(7)
FP64, FP32 = 0, 1
SUM_FP64, SUM_FP32_KAHAN = 0, 1
HOST, DEVICE_M, DEVICE = 0, 1, 3

EXPORTED = [
    "g4acc_create", "g4acc_destroy", "g4acc_reset", "g4acc_accumulate", "g4acc_accumulate_batch",
    "g4acc_finalize", "g4acc_sync", "g4acc_get_g4", "g4acc_get_g4_normalized", "g4acc_sum_to", "g4acc_g4_size", "g4acc_bytes",
    "g4acc_stream", "g4acc_last_error", "g4acc_compute_G", "g4acc_g_size", "g4acc_accumulate_G",
    "g4acc_set_timing", "g4acc_get_timing", "g4acc_set_tuning", "g4acc_accumulate_queued",
    "g4acc_accumulate_single",
    "g4acc_flush", "g4acc_sync_copy", "g4acc_accumulate_batch_n", "g4acc_compute_m",
    "g4acc_sp_enable", "g4acc_sp_get_m_r_w", "g4acc_sp_size", "g4acc_compute_G_noalloc",
    "g4acc_create_model", "g4acc_accumulate_general", "g4acc_compute_G_general",
]


class LogicError(RuntimeError):
    """The reference's std::logic_error."""


class Geometry_c(ctypes.Structure):
    _fields_ = [
        ("n_bands", ctypes.c_int), ("n_sites", ctypes.c_int), ("beta", ctypes.c_double),
        ("n_freq", ctypes.c_int),
        ("n_w_ex", ctypes.c_int), ("w_ex", ctypes.POINTER(ctypes.c_int)),
        ("n_k_ex", ctypes.c_int), ("k_ex", ctypes.POINTER(ctypes.c_int)),
        ("n_channels", ctypes.c_int), ("channels", ctypes.POINTER(ctypes.c_int)),
        ("k_add", ctypes.POINTER(ctypes.c_int)), ("k_sub", ctypes.POINTER(ctypes.c_int)),
        ("dft", ctypes.c_void_p), ("band_phase", ctypes.c_void_p), ("g0", ctypes.c_void_p),
    ]


class Options_c(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("sum_mode", ctypes.c_int),
                ("max_walkers", ctypes.c_int), ("max_vertices", ctypes.c_int),
                ("device", ctypes.c_int)]


class Walker_c(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 2), ("m", ctypes.c_void_p * 2), ("ld", ctypes.c_int * 2),
                ("tau", ctypes.c_void_p * 2), ("site", ctypes.c_void_p * 2),
                ("band", ctypes.c_void_p * 2), ("sign", ctypes.c_double)]


class Model_c(ctypes.Structure):
    _fields_ = [("spin_symmetric", ctypes.c_int), ("complex_m", ctypes.c_int),
                ("non_density_density", ctypes.c_int)]


class WalkerGeneral_c(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 2), ("m", ctypes.c_void_p * 2), ("ld", ctypes.c_int * 2),
                ("tau", ctypes.c_void_p * 2), ("site", ctypes.c_void_p * 2),
                ("band", ctypes.c_void_p * 2), ("site_l", ctypes.c_void_p * 2),
                ("band_l", ctypes.c_void_p * 2), ("sign", ctypes.c_double * 2)]


_lib = None


def _share_hip_runtime():
    """Keeps one HIP runtime per process when torch is used beside the library.

    The torch wheel bundles its own libamdhip64.so / libhsa-runtime64.so (sonames .so.7 / .so.1)
    and its libraries ask for them by the unversioned names. Loaded after torch, libg4acc.so binds
    to torch's copy through the shared soname. Loaded first, it would bring in the system runtime
    under the versioned name only, and a later `import torch` would load a second runtime into
    the process (whose device enumeration then fails). Loading the runtime under the unversioned
    names first makes a later torch bind to the same one.
    """
    if "torch" in sys.modules:
        return
    for name in ("libhsa-runtime64.so", "libamdhip64.so"):
        try:
            ctypes.CDLL(name, mode=ctypes.RTLD_GLOBAL)
        except OSError:
            pass  # libg4acc.so then resolves the runtime by its own dependencies


def load_library(path: str = LIB_PATH):
    """Loads libg4acc.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with dca_amd/build_lib.sh "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    _share_hip_runtime()
    L = ctypes.CDLL(path)
    vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    sig = {
        "g4acc_create": (i32, [ctypes.POINTER(Geometry_c), ctypes.POINTER(Options_c), ctypes.POINTER(vp)]),
        "g4acc_create_model": (i32, [ctypes.POINTER(Geometry_c), ctypes.POINTER(Options_c),
                                     ctypes.POINTER(Model_c), ctypes.POINTER(vp)]),
        "g4acc_accumulate_general": (i32, [vp, ctypes.POINTER(WalkerGeneral_c), i32, i32]),
        "g4acc_compute_G_general": (i32, [vp, ctypes.POINTER(WalkerGeneral_c), i32, vp, sz]),
        "g4acc_destroy": (i32, [vp]),
        "g4acc_reset": (i32, [vp]),
        "g4acc_accumulate": (i32, [vp, ctypes.POINTER(Walker_c), i32]),
        "g4acc_accumulate_batch": (i32, [vp, ctypes.POINTER(Walker_c), i32, i32]),
        "g4acc_finalize": (i32, [vp]),
        "g4acc_sync": (i32, [vp]),
        "g4acc_get_g4": (i32, [vp, i32, vp, sz]),
        "g4acc_get_g4_normalized": (i32, [vp, i32, ctypes.c_double, vp, sz]),
        "g4acc_sp_enable": (i32, [vp, i32, ctypes.POINTER(ctypes.c_int), i32]),
        "g4acc_sp_get_m_r_w": (i32, [vp, i32, vp, sz]),
        "g4acc_sp_size": (sz, [vp]),
        "g4acc_sum_to": (i32, [vp, vp]),
        "g4acc_g4_size": (sz, [vp]),
        "g4acc_bytes": (sz, [vp]),
        "g4acc_stream": (vp, [vp]),
        "g4acc_last_error": (ctypes.c_char_p, [vp]),
        "g4acc_compute_G": (i32, [vp, ctypes.POINTER(Walker_c), i32, vp, sz]),
        "g4acc_compute_G_noalloc": (i32, [vp, ctypes.POINTER(Walker_c), vp, sz]),
        "g4acc_g_size": (sz, [vp]),
        "g4acc_accumulate_G": (i32, [vp, vp, ctypes.c_double]),
        "g4acc_set_timing": (i32, [vp, i32]),
        "g4acc_set_tuning": (i32, [vp, ctypes.c_char_p, i32]),
        "g4acc_accumulate_queued": (i32, [vp, ctypes.POINTER(Walker_c), i32]),
        "g4acc_accumulate_single": (i32, [vp, ctypes.POINTER(Walker_c), i32]),
        "g4acc_flush": (i32, [vp]),
        "g4acc_sync_copy": (i32, [vp]),
        "g4acc_accumulate_batch_n": (i32, [vp, ctypes.POINTER(Walker_c), ctypes.POINTER(vp), i32, i32]),
        "g4acc_compute_m": (i32, [vp, i32, vp, vp, i32, vp, i32]),
        "g4acc_get_timing": (i32, [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_longlong)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _raise(code: int, msg: str):
    if code == OUT_OF_MEMORY:
        raise MemoryError(msg)
    if code == DEVICE_ERROR:
        raise RuntimeError(msg)
    raise LogicError(msg)


class _Keep:
    """Holds numpy buffers alive while the library may read them."""

    def __init__(self):
        self.items = []

    def add(self, a):
        self.items.append(a)
        return a


def walker_struct(meas, keep: _Keep) -> Walker_c:
    """Host-memory walker from a measurement (M_pair, configs, sign)."""
    w = Walker_c()
    for s in range(2):
        n = int(meas.n[s])
        w.n[s] = n
        w.ld[s] = max(n, 1)
        if n:
            m = keep.add(np.asfortranarray(meas.M[s], np.float64))
            w.ld[s] = m.shape[0] if m.ndim == 2 else n
            t = keep.add(np.ascontiguousarray(meas.tau[s], np.float64))
            r = keep.add(np.ascontiguousarray(meas.site[s], np.int32))
            b = keep.add(np.ascontiguousarray(meas.band[s], np.int32))
            w.m[s], w.tau[s], w.site[s], w.band[s] = (m.ctypes.data, t.ctypes.data, r.ctypes.data,
                                                      b.ctypes.data)
    w.sign = float(meas.sign)
    return w


def is_general(meas) -> bool:
    """A measurement of the general model: complex M or sign, or vertices with left labels."""
    return (getattr(meas, "site_l", None) is not None or any(np.iscomplexobj(x) for x in meas.M)
            or isinstance(meas.sign, complex))


def general_walker_struct(meas, keep: _Keep, complex_m: bool) -> WalkerGeneral_c:
    """Host-memory walker of the general model (complex M as interleaved doubles when complex_m)."""
    w = WalkerGeneral_c()
    nd = getattr(meas, "site_l", None) is not None
    for s in range(2):
        n = int(meas.n[s])
        w.n[s] = n
        w.ld[s] = max(n, 1)
        if n:
            if not complex_m and np.iscomplexobj(meas.M[s]):
                raise LogicError("complex M needs a handle with complex_m")
            m = keep.add(np.asfortranarray(meas.M[s], np.complex128 if complex_m else np.float64))
            w.ld[s] = m.shape[0]
            t = keep.add(np.ascontiguousarray(meas.tau[s], np.float64))
            r = keep.add(np.ascontiguousarray(meas.site[s], np.int32))
            b = keep.add(np.ascontiguousarray(meas.band[s], np.int32))
            w.m[s], w.tau[s], w.site[s], w.band[s] = (m.ctypes.data, t.ctypes.data, r.ctypes.data,
                                                      b.ctypes.data)
            if nd:
                rl = keep.add(np.ascontiguousarray(meas.site_l[s], np.int32))
                bl = keep.add(np.ascontiguousarray(meas.band_l[s], np.int32))
                w.site_l[s], w.band_l[s] = rl.ctypes.data, bl.ctypes.data
    sign = complex(meas.sign)
    w.sign[0], w.sign[1] = sign.real, sign.imag
    return w


def device_walker_struct(n: Sequence[int], m_ptr, ld, tau_ptr, site_ptr, band_ptr, sign) -> Walker_c:
    """Device-memory walker from raw device pointers (per spin)."""
    w = Walker_c()
    for s in range(2):
        w.n[s] = int(n[s])
        w.m[s] = int(m_ptr[s]) if n[s] else None
        w.ld[s] = int(ld[s])
        w.tau[s] = int(tau_ptr[s]) if n[s] else None
        w.site[s] = int(site_ptr[s]) if n[s] else None
        w.band[s] = int(band_ptr[s]) if n[s] else None
    w.sign = float(sign)
    return w


class TpAccumulator:
    """g4acc-backed TpAccumulator (one HIP device, one stream)."""

    def __init__(self, geom: geo.Geometry, precision: str = "fp64", sum_mode: str = "fp64",
                 max_walkers: int = 1, max_vertices: int = 64, device: int = 0, tuned: bool = True,
                 spin_symmetric: bool = True, complex_m: bool = False, non_density_density: bool = False):
        self.lib = load_library()
        self.geom = geom
        self._keep = _Keep()
        k = self._keep
        wex = k.add(np.ascontiguousarray(geom.w_ex, np.int32))
        kex = k.add(np.ascontiguousarray(geom.k_ex, np.int32))
        ch = k.add(np.ascontiguousarray(geom.channels, np.int32))
        add = k.add(np.ascontiguousarray(geom.cluster.add.ravel(order="F"), np.int32))
        sub = k.add(np.ascontiguousarray(geom.cluster.sub.ravel(order="F"), np.int32))
        dft = k.add(np.ascontiguousarray(geom.cluster.dft.ravel(order="F"), np.complex128))
        g0 = k.add(geo.g0_flat(geom.g0))
        bp = None
        if geom.band_phase is not None:
            bp = k.add(np.ascontiguousarray(geom.band_phase.ravel(order="F"), np.complex128))
        gc = Geometry_c()
        gc.n_bands, gc.n_sites, gc.beta, gc.n_freq = geom.nb, geom.nc, geom.beta, geom.n_freq
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        gc.n_w_ex, gc.w_ex = len(wex), ip(wex)
        gc.n_k_ex, gc.k_ex = len(kex), ip(kex)
        gc.n_channels, gc.channels = len(ch), ip(ch)
        gc.k_add, gc.k_sub = ip(add), ip(sub)
        gc.dft, gc.g0 = dft.ctypes.data, g0.ctypes.data
        gc.band_phase = None if bp is None else bp.ctypes.data
        opt = Options_c()
        opt.precision = {"fp64": FP64, "fp32": FP32}[precision]
        opt.sum_mode = {"fp64": SUM_FP64, "fp32_kahan": SUM_FP32_KAHAN}[sum_mode]
        opt.max_walkers, opt.max_vertices, opt.device = int(max_walkers), int(max_vertices), int(device)
        self.precision, self.sum_mode = precision, sum_mode
        self.max_walkers, self.max_vertices = int(max_walkers), int(max_vertices)
        h = ctypes.c_void_p()
        model = Model_c(int(spin_symmetric), int(complex_m), int(non_density_density))
        self.general_model = not spin_symmetric or complex_m or non_density_density
        self.complex_m = complex_m
        rc = self.lib.g4acc_create_model(ctypes.byref(gc), ctypes.byref(opt), ctypes.byref(model), ctypes.byref(h))
        if rc != OK:
            _raise(rc, self.lib.g4acc_last_error(None).decode())
        self._h = h
        self._inflight = []
        self.tuning = None
        if tuned and not self.general_model:
            # Launch shapes measured by dca_amd/autotune.py for this geometry and batch size.
            from . import autotune
            self.tuning = autotune.lookup(geom, precision, self.max_walkers)
            if self.tuning:
                self.set_tuning(**self.tuning)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.g4acc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != OK:
            _raise(rc, self.lib.g4acc_last_error(self._h).decode())

    # -- reference surface ---------------------------------------------------------------------
    def resetAccumulation(self, dca_loop: int = 0):
        self._check(self.lib.g4acc_reset(self._h))

    def accumulate(self, meas) -> None:
        """accumulate(M_pair, configs, sign) for one measurement in host memory."""
        if self.general_model or is_general(meas):
            return self.accumulate_batch([meas])
        keep = _Keep()
        w = walker_struct(meas, keep)
        self._check(self.lib.g4acc_accumulate(self._h, ctypes.byref(w), HOST))
        self._inflight.append(keep)

    def accumulate_batch(self, measurements) -> None:
        keep = _Keep()
        if self.general_model or any(is_general(m) for m in measurements):
            arr = (WalkerGeneral_c * len(measurements))(
                *[general_walker_struct(m, keep, self.complex_m) for m in measurements])
            self._check(self.lib.g4acc_accumulate_general(self._h, arr, len(measurements), HOST))
        else:
            arr = (Walker_c * len(measurements))(*[walker_struct(m, keep) for m in measurements])
            self._check(self.lib.g4acc_accumulate_batch(self._h, arr, len(measurements), HOST))
        self._inflight.append(keep)

    def accumulate_batch_n(self, measurements, exp_v) -> None:
        """Batch of host measurements whose `M` fields hold the walkers' N matrices, with
        exp_v[i][s] the exp(V) values of their vertices: M = diag(exp_v - 1) N is formed on the
        device (CtauxWalker::computeM, ctaux_walker.hpp:1889-1916)."""
        keep = _Keep()
        arr = (Walker_c * len(measurements))(*[walker_struct(m, keep) for m in measurements])
        xv = (ctypes.c_void_p * (2 * len(measurements)))()
        for i, pair in enumerate(exp_v):
            for s in range(2):
                a = keep.add(np.ascontiguousarray(pair[s], np.float64))
                xv[2 * i + s] = a.ctypes.data if a.size else None
        self._check(self.lib.g4acc_accumulate_batch_n(self._h, arr, xv, len(measurements), HOST))
        self._inflight.append(keep)

    def compute_m(self, n: int, exp_v_ptr: int, n_ptr: int, ldn: int, m_ptr: int, ldm: int) -> None:
        """M = diag(exp_v - 1) N on the device for device pointers (CtauxWalker::computeM), on the
        handle's stream."""
        self._check(self.lib.g4acc_compute_m(self._h, int(n), exp_v_ptr, n_ptr, int(ldn), m_ptr, int(ldm)))

    def accumulate_queued(self, meas, memory: int = HOST, walker: Walker_c | None = None) -> None:
        """accumulate(M_pair, configs, sign) through the handle's queue: the measurement is copied
        and runs in a batch of max_walkers (or at flush / sync / get_G4 / sumTo). Host
        measurements are given as `meas`; device ones as a ready `walker` (device_walker_struct)."""
        if walker is None:
            keep = _Keep()
            walker = walker_struct(meas, keep)
        self._check(self.lib.g4acc_accumulate_queued(self._h, ctypes.byref(walker), memory))

    def flush(self) -> None:
        self._check(self.lib.g4acc_flush(self._h))

    def sync_copy(self) -> None:
        """synchronizeCopy for queued measurements: their inputs have been copied."""
        self._check(self.lib.g4acc_sync_copy(self._h))

    def accumulate_device(self, walkers: Sequence[Walker_c]) -> None:
        """Batch of walkers whose arrays are device memory (see device_walker_struct)."""
        arr = (Walker_c * len(walkers))(*walkers)
        self._check(self.lib.g4acc_accumulate_batch(self._h, arr, len(walkers), DEVICE))

    def synchronizeCopy(self):
        self._check(self.lib.g4acc_sync(self._h))
        self._inflight.clear()

    def finalize(self):
        self._check(self.lib.g4acc_finalize(self._h))
        self._inflight.clear()

    def get_G4(self) -> list[np.ndarray]:
        """One flat complex128 vector per channel in the reference order
        (b1, b2, b3, b4, k1, w1, k2, w2, k_ex, w_ex), first index fastest."""
        n = self.lib.g4acc_g4_size(self._h)
        res = []
        for c in range(len(self.geom.channels)):
            out = np.empty(n, np.complex128)
            self._check(self.lib.g4acc_get_g4(self._h, c, out.ctypes.data, n))
            res.append(out)
        self._inflight.clear()
        return res

    def get_G4_normalized(self, accumulated_sign: float) -> list[np.ndarray]:
        """get_G4 divided by beta^2 * accumulated_sign, the solvers' normalisation
        (ctaux_cluster_solver.hpp:523-527); accumulated_sign is the sum of the measurement signs."""
        n = self.lib.g4acc_g4_size(self._h)
        res = []
        for c in range(len(self.geom.channels)):
            out = np.empty(n, np.complex128)
            self._check(self.lib.g4acc_get_g4_normalized(self._h, c, float(accumulated_sign),
                                                         out.ctypes.data, n))
            res.append(out)
        self._inflight.clear()
        return res

    def write_G4_hdf5(self, path: str, accumulated_sign: float | None = None):
        """Writes every channel's G4 (normalised by accumulated_sign beta^2 when given) to `path`
        in the reference's HDF5 format: /functions/G4_<channel> (dca_data.hpp:331-336,
        hdf5_writer.hpp:244-262), through libg4acc_io.so."""
        from . import g4io
        g4 = self.get_G4() if accumulated_sign is None else self.get_G4_normalized(accumulated_sign)
        g4io.write_g4(path, self.geom, g4)

    def enable_sp(self, n_sp_freq: int, accumulate_m_sqr: bool = False):
        """Adds the single-particle accumulator (SpAccumulator, sp_accumulator_gpu.hpp) to this
        handle: every accumulate then also adds the measurement's M_r_w."""
        rs = np.ascontiguousarray(self.geom.cluster.rsub.ravel(order="F"), np.int32)
        self._check(self.lib.g4acc_sp_enable(self._h, int(n_sp_freq),
                                             rs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                             1 if accumulate_m_sqr else 0))
        self.n_sp_freq = int(n_sp_freq)

    def get_sign_times_M_r_w(self, squared: bool = False) -> np.ndarray:
        """M_r_w (or M_r_w_sqr) shaped (b1, s1, b2, s2, r, w) (Fortran order)."""
        n = self.lib.g4acc_sp_size(self._h)
        out = np.empty(n, np.complex128)
        self._check(self.lib.g4acc_sp_get_m_r_w(self._h, 1 if squared else 0, out.ctypes.data, n))
        g = self.geom
        return out.reshape((g.nb, 2, g.nb, 2, g.nc, 2 * self.n_sp_freq), order="F")

    def sumTo(self, other: "TpAccumulator"):
        self._check(self.lib.g4acc_sum_to(self._h, other._h))

    # -- testing / profiling hooks -----------------------------------------------------------------
    def compute_G(self, meas) -> np.ndarray:
        """G of one measurement in the reference layout (b1, b2, s, k1, k2, w1E, w2E)."""
        keep = _Keep()
        n = self.lib.g4acc_g_size(self._h)
        out = np.empty(n, np.complex128)
        if self.general_model or is_general(meas):
            w = general_walker_struct(meas, keep, self.complex_m)
            self._check(self.lib.g4acc_compute_G_general(self._h, ctypes.byref(w), HOST, out.ctypes.data, n))
        else:
            w = walker_struct(meas, keep)
            self._check(self.lib.g4acc_compute_G(self._h, ctypes.byref(w), HOST, out.ctypes.data, n))
        g = self.geom
        return out.reshape((g.nb, g.nb, 2, g.nc, g.nc, g.n_ext, g.n_ext), order="F")

    def accumulate_G(self, G: np.ndarray, sign: float):
        buf = np.ascontiguousarray(np.asarray(G, np.complex128).ravel(order="F"))
        self._check(self.lib.g4acc_accumulate_G(self._h, buf.ctypes.data, float(sign)))

    @property
    def stream(self) -> int:
        return self.lib.g4acc_stream(self._h) or 0

    @property
    def device_bytes(self) -> int:
        return self.lib.g4acc_bytes(self._h)

    def set_tuning(self, **params):
        """Launch-shape tuning: mtog_tw, mtog_mfma, g4_qg, g4_tile_order, g4_mfma, g4_rank1, graph
        (include/g4acc.h)."""
        for k, v in params.items():
            self._check(self.lib.g4acc_set_tuning(self._h, k.encode(), int(v)))

    def set_timing(self, enable: bool):
        self._check(self.lib.g4acc_set_timing(self._h, 1 if enable else 0))

    def get_timing(self):
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        self._check(self.lib.g4acc_get_timing(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value
