"""The kernels' clipping code (voronota-ltr_amd/csrc/contour.h and
group_clip.h), built for the host, against the oracle pair by pair: every
record bit-identical, through the one-lane path (the large-contour kernel),
the classification + masked path, and the 16-lane group path of the fused
clip kernel run on 16 lock-stepped host fibers. Also the block remap used by
every wavefront-per-ball kernel."""
import ctypes as C

import numpy as np
import pytest

import hostclip
from conftest import RING17, corners_box, load_xyzr


def _staged(sph4, off, idx, pairs, mode=0):
    L = hostclip._load()
    f = L.clip_host_records_staged
    f.argtypes = [C.c_void_p] * 4 + [C.c_int64, C.c_int, C.c_void_p, C.c_void_p]
    sph4 = np.ascontiguousarray(sph4, np.float64)
    off = np.ascontiguousarray(off, np.int64)
    idx = np.ascontiguousarray(idx, np.int32)
    pairs = np.ascontiguousarray(pairs, np.int32)
    out = np.zeros((len(pairs), 6))
    fl = np.zeros(len(pairs), np.uint32)
    f(sph4.ctypes.data, off.ctypes.data, idx.ctypes.data, pairs.ctypes.data, len(pairs), mode, out.ctypes.data,
      fl.ctypes.data)
    return out, fl


def _group(sph4, off, idx, pairs, mode=0):
    L = hostclip._load()
    f = L.clip_host_records_group
    f.argtypes = [C.c_void_p] * 4 + [C.c_int64, C.c_int, C.c_void_p, C.c_void_p]
    sph4 = np.ascontiguousarray(sph4, np.float64)
    off = np.ascontiguousarray(off, np.int64)
    idx = np.ascontiguousarray(idx, np.int32)
    pairs = np.ascontiguousarray(pairs, np.int32)
    out = np.zeros((len(pairs), 6))
    fl = np.zeros(len(pairs), np.uint32)
    f(sph4.ctypes.data, off.ctypes.data, idx.ctypes.data, pairs.ctypes.data, len(pairs), mode, out.ctypes.data,
      fl.ctypes.data)
    return out, fl


CASES = [
    ("cs_1x1", lambda: load_xyzr("balls_cs_1x1.xyzr"), 2.0, None),
    ("1ctf", lambda: load_xyzr("1ctf_pdb.xyzr"), 1.4, None),
    ("2zsk", lambda: load_xyzr("balls_2zsk.xyzr"), 1.4, None),
    ("ring17_periodic", lambda: RING17, 1.0, corners_box((-1.6, -1.6, -0.6), (1.6, 1.6, 3.1))),
    ("cs_1x1_periodic", lambda: load_xyzr("balls_cs_1x1.xyzr"), 2.0, corners_box((0, 0, 0), (200, 250, 300))),
]


@pytest.mark.parametrize("name,make,probe,pbox", CASES)
@pytest.mark.parametrize("mode", [0, 1])
def test_records_bit_identical(oracle, name, make, probe, pbox, mode):
    balls = make()
    off, idx, _, _, pairs = oracle.collisions(balls, probe, pbox=pbox)
    ref, rf = oracle.pair_records(balls, probe, len(pairs), pbox=pbox, mode=mode)
    sph4 = hostclip.populated(balls, probe, pbox)
    got, gf = hostclip.records(sph4, off, idx, pairs, mode=mode)
    assert np.array_equal(gf & 3, rf & 3)
    assert np.array_equal(got, ref)
    got2, gf2 = _staged(sph4, off, idx, pairs, mode=mode)
    assert np.array_equal(gf2 & 3, rf & 3)
    assert np.array_equal(got2, ref)
    if name != "2zsk":  # (82k pairs on fibers take ~40 s; 1ctf and RSA cover proteins)
        got3, gf3 = _group(sph4, off, idx, pairs, mode=mode)
        assert not (gf3 & 64).any(), "lanes of a group disagreed"
        assert np.array_equal(gf3 & 3, rf & 3)
        assert np.array_equal(got3, ref)
        assert (gf3 & 16).any()


def test_records_bit_identical_rsa(oracle):
    from voronota_ltr import workload

    balls, _ = workload.rsa_sphere(2500, seed=2500)
    off, idx, _, _, pairs = oracle.collisions(balls, 1.4)
    ref, rf = oracle.pair_records(balls, 1.4, len(pairs))
    got, gf = _staged(hostclip.populated(balls, 1.4), off, idx, pairs)
    assert np.array_equal(gf & 3, rf & 3) and np.array_equal(got, ref)
    # the prefilter alone settles a large share of the pairs
    assert (gf & 8).sum() / 8 > 0.3 * len(pairs)
    got3, gf3 = _group(hostclip.populated(balls, 1.4), off, idx, pairs)
    assert not (gf3 & 64).any()
    assert np.array_equal(gf3 & 3, rf & 3) and np.array_equal(got3, ref)
    # nearly every cut pair stays within the group's 16 / 32 points
    assert ((gf3 & 32) > 0).sum() < 1e-3 * ((gf3 & 16) > 0).sum()


def test_xcd_remap_is_bijective():
    L = hostclip._load()
    L.xcd_remap_host.restype = C.c_int64
    L.xcd_remap_host.argtypes = [C.c_int64, C.c_int64]
    for g in list(range(1, 130)) + [1000, 4097, 25000]:
        assert sorted(L.xcd_remap_host(b, g) for b in range(g)) == list(range(g))
