"""ctypes binding of libg4acc_io.so (include/g4acc_io.h): G4 in the reference's HDF5 format.

Writing follows HDF5Writer::execute for a complex function (include/dca/io/hdf5/hdf5_writer.hpp:
244-262): group "functions", dataset "G4_<channel>" (dca_data.hpp:331-336), compound {r, i}
doubles, dimensions stored slowest first, attributes "domain-sizes" and "name".
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libg4acc_io.so")
_lib = None


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with dca_amd/build_lib.sh")
    L = ctypes.CDLL(path)
    u64p = ctypes.POINTER(ctypes.c_ulonglong)
    L.g4io_write_function.restype = ctypes.c_int
    L.g4io_write_function.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p,
                                      u64p, ctypes.c_int, ctypes.c_int]
    L.g4io_read_function.restype = ctypes.c_int
    L.g4io_read_function.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t,
                                     u64p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.g4io_read_name.restype = ctypes.c_int
    L.g4io_read_name.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
    L.g4io_channel_name.restype = ctypes.c_char_p
    L.g4io_channel_name.argtypes = [ctypes.c_int]
    L.g4io_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


class IOError_(OSError):
    pass


def _check(rc):
    if rc != 0:
        raise IOError_(load_library().g4io_last_error().decode())


def channel_name(channel: int) -> str:
    n = load_library().g4io_channel_name(int(channel))
    if n is None:
        raise ValueError(f"unknown channel {channel}")
    return n.decode()


def write_function(path: str, group: str, name: str, values: np.ndarray, dims, truncate: bool = False):
    """Writes a complex function (values flat, first index fastest) with dimensions `dims`."""
    v = np.ascontiguousarray(np.asarray(values, np.complex128).ravel())
    d = (ctypes.c_ulonglong * len(dims))(*[int(x) for x in dims])
    if int(np.prod(dims)) != v.size:
        raise ValueError("dims do not match the number of values")
    _check(load_library().g4io_write_function(path.encode(), group.encode(), name.encode(), v.ctypes.data, d,
                                              len(dims), 1 if truncate else 0))


def read_function(path: str, dataset: str):
    """Returns (values flat first index fastest, dims first index fastest)."""
    L = load_library()
    d = (ctypes.c_ulonglong * 32)()
    n = ctypes.c_int()
    _check(L.g4io_read_function(path.encode(), dataset.encode(), None, 0, d, 32, ctypes.byref(n)))
    dims = tuple(int(d[k]) for k in range(n.value))
    out = np.empty(int(np.prod(dims)), np.complex128)
    _check(L.g4io_read_function(path.encode(), dataset.encode(), out.ctypes.data, out.size, d, 32, ctypes.byref(n)))
    return out, dims


def read_name(path: str, dataset: str) -> str:
    buf = ctypes.create_string_buffer(256)
    _check(load_library().g4io_read_name(path.encode(), dataset.encode(), buf, 256))
    return buf.value.decode()


def write_g4(path: str, geometry, g4_per_channel, truncate: bool = True):
    """All channels of a G4 (get_G4 order) as /functions/G4_<channel> (dca_data.hpp:331-336)."""
    for i, (ch, g4) in enumerate(zip(geometry.channels, g4_per_channel)):
        write_function(path, "functions", "G4_" + channel_name(ch), g4, geometry.g4_shape(),
                       truncate=truncate and i == 0)
