"""Seeded synthetic volumes of the benchmark workloads (native generator in
libxd_synth.so, csrc/xd_synth.cpp).

  shells_128()  128^3: 24 ellipsoidal emission shells (semi-axes 8-40 voxels,
                walls 2-4 voxels) with uniform random colours plus 5 % sparse
                noise voxels.
  ct_body()     512x361x512 CT-scan-like body: Gaussian-smoothed (sigma 6)
                random blobs thresholded into a closed thick wall with an
                interior density gradient; density -> colour by a linear ramp;
                about 70 % of voxels empty.
  lognormal()   cosmological-gas-like volume (1024^3 or 2048^3): a Gaussian
                random field with P(k) ~ k^-3 on a 256^3 grid (in-repo FFT),
                exponentiated (log-normal), upsampled trilinearly, voxels below
                the 90th percentile zeroed; emission = log-density through a
                colour ramp.

Arrays are [Z, Y, X, 4] uint8 (x fastest), the layout of the reference's Grid.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libxd_synth.so")

SEED_SHELLS = 0x5EED0001C0FFEE01
SEED_CT = 0x5EED0002C0FFEE02
SEED_LOGNORMAL = 0x5EED0003C0FFEE03

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m xenodon_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.xds_hash.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.xds_hash.restype = ctypes.c_uint64
        L.xds_shells.argtypes = [ctypes.c_void_p] + [ctypes.c_uint64] * 4 + [ctypes.c_uint32] + [ctypes.c_double] * 5
        L.xds_ct_body.argtypes = [ctypes.c_void_p] + [ctypes.c_uint64] * 4 + [ctypes.c_uint32] + [ctypes.c_double] * 3
        L.xds_lognormal.argtypes = [ctypes.c_void_p] + [ctypes.c_uint64] * 4 + [ctypes.c_uint32] + \
            [ctypes.c_double] * 3
        _lib = L
    return _lib


def volume_hash(arr):
    a = np.ascontiguousarray(arr)
    return int(lib().xds_hash(a.ctypes.data, a.nbytes))


def shells(shape_xyz=(128, 128, 128), seed=SEED_SHELLS, n_shells=24, radii=(8.0, 40.0), thickness=(2.0, 4.0),
           noise=0.05):
    X, Y, Z = shape_xyz
    out = np.empty((Z, Y, X, 4), np.uint8)
    lib().xds_shells(out.ctypes.data, X, Y, Z, seed, n_shells, radii[0], radii[1], thickness[0], thickness[1],
                     noise)
    return out


def shells_128():
    return shells()


def ct_body(shape_xyz=(512, 361, 512), seed=SEED_CT, n_blobs=40, sigma=6.0, empty=0.70, wall=0.35):
    X, Y, Z = shape_xyz
    out = np.empty((Z, Y, X, 4), np.uint8)
    lib().xds_ct_body(out.ctypes.data, X, Y, Z, seed, n_blobs, sigma, empty, wall)
    return out


def lognormal(side=1024, seed=SEED_LOGNORMAL, base=256, slope=-3.0, sigma_ln=1.0, keep_above=0.90):
    """side^3 log-normal volume ([Z, Y, X, 4] uint8); 1024 -> 4 GiB, 2048 -> 32 GiB."""
    out = np.empty((side, side, side, 4), np.uint8)
    lib().xds_lognormal(out.ctypes.data, side, side, side, seed, base, slope, sigma_ln, keep_above)
    return out
