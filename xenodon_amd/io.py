"""ctypes binding of the xenodon CLI's file formats (libxd_io.so): TIFF stacks
(read with TIFFReadRGBAImage semantics, write), .svo octrees and PNG frames.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libxd_io.so")
CLI_PATH = os.path.join(PKG, "bin", "xenodon")

_lib = None


class IoError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m xenodon_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        P, vp, u64 = ctypes.POINTER, ctypes.c_void_p, ctypes.c_uint64
        sigs = {
            "xdio_load_tiff": [ctypes.c_char_p, P(u64), P(vp), ctypes.c_char_p, ctypes.c_size_t],
            "xdio_save_tiff": [ctypes.c_char_p, P(u64), vp, ctypes.c_char_p, ctypes.c_size_t],
            "xdio_load_svo": [ctypes.c_char_p, P(u64), P(u64), P(vp), ctypes.c_char_p, ctypes.c_size_t],
            "xdio_save_svo": [ctypes.c_char_p, u64, u64, vp, ctypes.c_char_p, ctypes.c_size_t],
            "xdio_save_png": [ctypes.c_char_p, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t],
            "xdio_format_output_path": [ctypes.c_char_p, u64, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p,
                                        ctypes.c_size_t],
        }
        for name, args in sigs.items():
            getattr(L, name).argtypes = args
            getattr(L, name).restype = ctypes.c_int
        L.xdio_free.argtypes = [vp]
        L.xdio_free.restype = None
        _lib = L
    return _lib


def _call(fn, *args):
    err = ctypes.create_string_buffer(512)
    if fn(*args, err, len(err)) != 0:
        raise IoError(err.value.decode())


def load_tiff(path):
    """Grid as uint8 [Z, Y, X, 4] (RGBA), as Grid::load_tiff reads it."""
    dims = (ctypes.c_uint64 * 3)()
    p = ctypes.c_void_p()
    _call(lib().xdio_load_tiff, str(path).encode(), dims, ctypes.byref(p))
    X, Y, Z = dims
    try:
        n = X * Y * Z * 4
        out = np.frombuffer(ctypes.string_at(p, n), np.uint8).reshape(Z, Y, X, 4).copy()
    finally:
        lib().xdio_free(p)
    return out


def save_tiff(path, grid):
    g = np.ascontiguousarray(grid, np.uint8)
    dims = (ctypes.c_uint64 * 3)(g.shape[2], g.shape[1], g.shape[0])
    _call(lib().xdio_save_tiff, str(path).encode(), dims, g.ctypes.data)


def load_svo(path):
    """(nodes uint32 [N, 10], side)."""
    side, n = ctypes.c_uint64(), ctypes.c_uint64()
    p = ctypes.c_void_p()
    _call(lib().xdio_load_svo, str(path).encode(), ctypes.byref(side), ctypes.byref(n), ctypes.byref(p))
    try:
        words = np.frombuffer(ctypes.string_at(p, n.value * 40), np.uint32).reshape(-1, 10).copy()
    finally:
        lib().xdio_free(p)
    return words, side.value


def save_svo(path, nodes, side):
    n = np.ascontiguousarray(nodes, np.uint32).reshape(-1, 10)
    _call(lib().xdio_save_svo, str(path).encode(), int(side), n.shape[0], n.ctypes.data)


def save_png(path, rgba):
    img = np.ascontiguousarray(rgba, np.uint8)
    _call(lib().xdio_save_png, str(path).encode(), img.ctypes.data, img.shape[1], img.shape[0])


def format_output_path(fmt, frame):
    out = ctypes.create_string_buffer(4096)
    _call(lib().xdio_format_output_path, fmt.encode(), frame, out, len(out))
    return out.value.decode()
