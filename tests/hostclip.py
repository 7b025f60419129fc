"""Loader of tests/native/libclip_host.so: the kernels' clipping code built
for the host (test helper)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VLTR_CLIP_HOST_LIB: the sanitizer build (tests/native/_san/), loaded by
# tests/test_sanitizers.py
LIB = os.environ.get("VLTR_CLIP_HOST_LIB") or os.path.join(HERE, "native", "libclip_host.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB)
        L.clip_host_records.argtypes = [C.c_void_p] * 4 + [C.c_int64, C.c_int, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def records(sph4, nb_off, nb_idx, pairs, mode=0):
    """Per-pair (P x 6 values, P flags) from the host build of contour.h."""
    L = _load()
    sph4 = np.ascontiguousarray(sph4, np.float64)
    nb_off = np.ascontiguousarray(nb_off, np.int64)
    nb_idx = np.ascontiguousarray(nb_idx, np.int32)
    pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    P = len(pairs)
    out = np.zeros((P, 6))
    flags = np.zeros(P, np.uint32)
    L.clip_host_records(sph4.ctypes.data, nb_off.ctypes.data, nb_idx.ctypes.data, pairs.ctypes.data, P, mode,
                        out.ctypes.data, flags.ctypes.data)
    return out, flags


def populated(balls, probe, pbox=None):
    """(x, y, z, R) of the searched spheres: the balls, then 26 images per
    lattice shift in the reference's order when periodic (types.rs:362-375)."""
    b = np.asarray(balls, np.float64).reshape(-1, 4)
    s = b.copy()
    s[:, 3] = b[:, 3] + probe
    if pbox is None:
        return s
    v = np.asarray(pbox, np.float64).reshape(3, 3)
    out = [s]
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sz in (-1, 0, 1):
                if sx == sy == sz == 0:
                    continue
                t = s.copy()
                for k in range(3):
                    # fma chain c*wc + (b*wb + (a*wa + x)); weights are exact small ints
                    t[:, k] = ((s[:, k] + v[0, k] * sx) + v[1, k] * sy) + v[2, k] * sz
                out.append(t)
    return np.concatenate(out)
