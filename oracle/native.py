"""ctypes loader for the C oracle (oracle/c/oracle.c).  Test infrastructure only."""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)


def build(force=False):
    """Compile the C oracle with its Makefile (gcc, no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_project.argtypes = [_dp, _dp, ctypes.c_int64, _dp, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double, _dp, _i32p, _i32p]
        L.oracle_window_eval.restype = ctypes.c_double
        L.oracle_window_eval.argtypes = [ctypes.c_int] + [ctypes.c_double] * 5
        L.oracle_octant_rfft_kernel.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                                ctypes.c_double, _dp]
        L.oracle_full_rfft_kernel.argtypes = L.oracle_octant_rfft_kernel.argtypes
        L.oracle_spherical_jn.restype = ctypes.c_double
        L.oracle_spherical_jn.argtypes = [ctypes.c_int, ctypes.c_double]
        L.oracle_assoc_legendre.restype = ctypes.c_double
        L.oracle_assoc_legendre.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.oracle_sph_harm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, _dp, _dp]
        L.oracle_radial_multipole_window.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int,
                                                     ctypes.c_int, _dp]
        L.oracle_profile_multipole_window.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                                      ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_pair_half_multipole.restype = ctypes.c_double
        L.oracle_pair_half_multipole.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int, _dp]
        L.oracle_summand.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_interpolate.argtypes = [_dp, _dp, ctypes.c_int64, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.oracle_triangle_offsets.argtypes = [ctypes.c_double] * 6 + [_dp]
        L.oracle_triplet_box_random.restype = ctypes.c_double
        L.oracle_triplet_box_random.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                                ctypes.c_int64, ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_int]
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.oracle_mc_binned.argtypes = ([_dp] * 4 + [ctypes.c_int] * 5 + [ctypes.c_double] * 4 +
                                       [ctypes.c_int64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint8), _i32p,
                                        ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_float), _dp,
                                        ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                        ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                        ctypes.POINTER(ctypes.c_int64)])
        _lib = L
    return _lib


def ptr(a, ctype=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def num_threads():
    return int(lib().oracle_num_threads())
