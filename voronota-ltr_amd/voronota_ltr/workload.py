"""Seeded synthetic workloads (libvltr_workload.so, plain C++).

Random sequential addition of balls at 0.056 balls/A^3 with a 2.0 A minimum
centre distance and radii uniform from {1.52, 1.55, 1.70, 1.80} A; the radius
index is the ball's type. SplitMix64 streams make inputs identical on every
machine.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvltr_workload.so")
DENSITY = 0.056
MIN_DIST = 2.0
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"workload library not built: {LIB_PATH}")
        L = C.CDLL(LIB_PATH)
        f64p = C.POINTER(C.c_double)
        i32p = C.POINTER(C.c_int32)
        L.vltr_workload_rsa_sphere.argtypes = [C.c_uint64, C.c_size_t, C.c_double, C.c_double, f64p, i32p, C.c_uint64]
        L.vltr_workload_rsa_sphere.restype = C.c_size_t
        L.vltr_workload_rsa_box.argtypes = [C.c_uint64, C.c_size_t, C.c_double, C.c_double, f64p, i32p, C.c_uint64]
        L.vltr_workload_rsa_box.restype = C.c_size_t
        L.vltr_workload_batch_sizes.argtypes = [C.c_uint64, C.c_size_t, C.c_double, C.c_double,
                                                C.POINTER(C.c_uint64)]
        L.vltr_workload_type_matrix.argtypes = [C.c_uint64, f64p]
        L.vltr_workload_uniform.argtypes = [C.c_uint64, C.c_size_t, f64p]
        _lib = L
    return _lib


def sphere_radius(n: int, density: float = DENSITY) -> float:
    return (3.0 * n / (4.0 * math.pi * density)) ** (1.0 / 3.0)


def rsa_sphere(n: int, seed: int, density: float = DENSITY, min_dist: float = MIN_DIST):
    """(balls (n,4), types (n,)) inside a sphere holding n balls at `density`."""
    L = _load()
    balls = np.zeros((n, 4), dtype=np.float64)
    types = np.zeros(n, dtype=np.int32)
    got = L.vltr_workload_rsa_sphere(seed, n, density, min_dist, balls.ctypes.data_as(C.POINTER(C.c_double)),
                                     types.ctypes.data_as(C.POINTER(C.c_int32)), 1 << 40)
    if got != n:
        raise RuntimeError(f"RSA placed {got} of {n} balls")
    return balls, types


def rsa_box(n: int, edge: float, seed: int, min_dist: float = MIN_DIST):
    """(balls, types) in the periodic cube [0, edge)^3 (minimum-image distances)."""
    L = _load()
    balls = np.zeros((n, 4), dtype=np.float64)
    types = np.zeros(n, dtype=np.int32)
    got = L.vltr_workload_rsa_box(seed, n, edge, min_dist, balls.ctypes.data_as(C.POINTER(C.c_double)),
                                  types.ctypes.data_as(C.POINTER(C.c_int32)), 1 << 40)
    if got != n:
        raise RuntimeError(f"RSA placed {got} of {n} balls")
    return balls, types


def batch_sizes(n_structs: int, seed: int, lo: float = 200.0, hi: float = 20000.0) -> np.ndarray:
    L = _load()
    out = np.zeros(n_structs, dtype=np.uint64)
    L.vltr_workload_batch_sizes(seed, n_structs, lo, hi, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


def type_matrix(seed: int) -> np.ndarray:
    L = _load()
    w = np.zeros(16, dtype=np.float64)
    L.vltr_workload_type_matrix(seed, w.ctypes.data_as(C.POINTER(C.c_double)))
    return w.reshape(4, 4)


def uniform(n: int, seed: int) -> np.ndarray:
    L = _load()
    out = np.zeros(n, dtype=np.float64)
    L.vltr_workload_uniform(seed, n, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out
