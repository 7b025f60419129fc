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


def _group(sph4, off, idx, pairs, mode=0, group_size=16):
    L = hostclip._load()
    f = L.clip_host_records_group
    f.argtypes = [C.c_void_p] * 4 + [C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    sph4 = np.ascontiguousarray(sph4, np.float64)
    off = np.ascontiguousarray(off, np.int64)
    idx = np.ascontiguousarray(idx, np.int32)
    pairs = np.ascontiguousarray(pairs, np.int32)
    out = np.zeros((len(pairs), 6))
    fl = np.zeros(len(pairs), np.uint32)
    f(sph4.ctypes.data, off.ctypes.data, idx.ctypes.data, pairs.ctypes.data, len(pairs), mode, out.ctypes.data,
      fl.ctypes.data, group_size)
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
        for gs in (16, 8):
            got3, gf3 = _group(sph4, off, idx, pairs, mode=mode, group_size=gs)
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
    got4, gf4 = _group(hostclip.populated(balls, 1.4), off, idx, pairs, group_size=8)
    assert not (gf4 & 64).any()
    assert np.array_equal(gf4 & 3, rf & 3) and np.array_equal(got4, ref)
    # with 8 lanes about 1% of the cut pairs outgrow the group
    assert ((gf4 & 32) > 0).sum() < 0.03 * ((gf4 & 16) > 0).sum()


def test_xcd_remap_is_bijective():
    L = hostclip._load()
    L.xcd_remap_host.restype = C.c_int64
    L.xcd_remap_host.argtypes = [C.c_int64, C.c_int64]
    for g in list(range(1, 130)) + [1000, 4097, 25000]:
        assert sorted(L.xcd_remap_host(b, g) for b in range(g)) == list(range(g))


def _adjoint_host(balls, probe, types, W):
    import ctypes as C

    L = hostclip._load()
    f = L.adjoint_host
    f.restype = C.c_int64
    f.argtypes = [C.c_void_p, C.c_int64] + [C.c_void_p] * 3 + [C.c_int64] + [C.c_void_p] * 4
    return f


@pytest.mark.parametrize("n,seed", [(40, 21), (120, 5)])
def test_adjoint_matches_dual_numbers(oracle, n, seed):
    """The kernels' adjoint code (adjoint.h: topology from the group clip,
    vertex recomputation, hand-written VJPs) run on host fibers, against the
    oracle's forward-mode dual numbers for every coordinate."""
    import ctypes as C

    from voronota_ltr import workload

    balls, types = workload.rsa_sphere(n, seed=seed)
    W = workload.type_matrix(3)
    off, idx, _, _, pairs = oracle.collisions(balls, 1.4)
    S = np.ascontiguousarray(hostclip.populated(balls, 1.4))
    off = np.ascontiguousarray(off, np.int64)
    idx = np.ascontiguousarray(idx, np.int32)
    pr = np.ascontiguousarray(pairs, np.int32)
    ty = np.ascontiguousarray(types, np.int32)
    Wc = np.ascontiguousarray(W, np.float64)
    e = C.c_double()
    g = np.zeros((n, 3))
    f = _adjoint_host(balls, 1.4, types, W)
    missing = f(S.ctypes.data, n, off.ctypes.data, idx.ctypes.data, pr.ctypes.data, len(pr), ty.ctypes.data,
                Wc.ctypes.data, C.byref(e), g.ctypes.data)
    assert missing == 0
    assert e.value == pytest.approx(oracle.energy(balls, 1.4, types, W), rel=1e-12)
    for ball in range(n):
        for axis in range(3):
            _, de = oracle.energy_dual(balls, 1.4, types, W, ball, axis)
            assert abs(g[ball, axis] - de) <= 1e-7 * max(1.0, abs(de)), (ball, axis, g[ball, axis], de)


def test_single_precision_cut_distance_bound():
    """The device's cut-distance test (`quotient_reaches` in contour.h) lets a
    single-precision estimate of l / sqrt(t) decide outside a 1e-5 relative
    margin around the circle radius. Its error: l and t rounded to single, a
    reciprocal square root within 2 ulp (the hardware's is within 1), one
    single product. Over the domain where the estimate is used (t and l in
    [1e-30, 1e15], t <= 1) it stays within 5e-7 of the exact double quotient,
    so an estimate beyond the margin can never disagree with the exact test."""
    rng = np.random.default_rng(7)
    t = 10.0 ** rng.uniform(-30, 0, 200000)
    l = 10.0 ** rng.uniform(-30, 15, 200000)
    exact = l / np.sqrt(t)
    tf = t.astype(np.float32)
    rs = (1.0 / np.sqrt(tf.astype(np.float64))).astype(np.float32)
    worst = 0.0
    for ulps in (-2, 0, 2):
        r = rs
        for _ in range(abs(ulps)):
            r = np.nextafter(r, np.float32(np.inf) if ulps > 0 else np.float32(0))
        est = (l.astype(np.float32) * r).astype(np.float64)
        ok = np.isfinite(est)
        assert ok.all()
        worst = max(worst, float(np.max(np.abs(est / exact - 1.0))))
    assert worst < 5e-7
    # and so the decision outside the margin is the exact one
    radius = exact * (1.0 + rng.uniform(-3e-5, 3e-5, exact.size))
    decided = (est > radius * (1.0 + 1e-5)) | (est < radius * (1.0 - 1e-5))
    assert np.array_equal((est > radius * (1.0 + 1e-5))[decided], (exact >= radius)[decided])


@pytest.mark.parametrize("n,per,extra,ok", [
    (100000, 24, 0, True),            # the 100k workload's size
    (524287, 4096, 4095, True),       # exactly INT32_MAX pairs
    (524288, 4096, 0, False),         # 2^31 pairs
    (1100000, 4096, 0, False),        # 4.5e9 pairs: the 32-bit total wraps to 2.1e8
    (2097152, 2048, 0, False),        # exactly 2^32 pairs: the 32-bit total wraps to 0
])
def test_pair_total_guard(n, per, extra, ok):
    """The candidate-pair total is refused beyond INT32_MAX (pair indices are
    int32), also where the 32-bit scan total has wrapped to a small value.
    Synthetic per-ball counts (at most 4096 each, as a ball has at most 4096
    listed neighbors) through the scan's own total arithmetic."""
    import ctypes as C

    L = hostclip._load()
    f = L.scan_pair_guard
    f.restype = C.c_int64
    f.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_uint32)]
    counts = np.full(n, per, np.uint32)
    counts[-1] += extra
    t32 = C.c_uint32()
    got = f(counts.ctypes.data, n, C.byref(t32))
    exact = int(counts.astype(np.uint64).sum())
    assert t32.value == exact % (1 << 32)
    if ok:
        assert got == exact
    else:
        assert got == -1, f"total {exact} (32-bit {t32.value}) was accepted"


def test_canonical_id_without_division():
    """canon_mod (vltr_internal.h): c mod n for sphere ids c < 27 n from the
    double reciprocal of n, truncated to int32 and corrected by one either
    way, equals the integer remainder: every id for small n, and ids around
    every multiple of n (where the truncated quotient can be off by one) for
    n up to the 2^31 / 27 balls the library accepts."""
    import numpy as np

    def canon(c, n):
        inv = np.float64(1.0) / np.float64(n)
        q = (c.astype(np.float64) * inv).astype(np.int64)  # truncation, as the int32 conversion
        r = c - q * n
        r = np.where(r < 0, r + n, np.where(r >= n, r - n, r))
        assert np.all(q <= 27) and np.all(q * n < 2**31)
        return r

    for n in range(1, 300):
        c = np.arange(27 * n, dtype=np.int64)
        assert np.array_equal(canon(c, n), c % n)
    rng = np.random.default_rng(27)
    cap = (1 << 31) // 27
    for n in [cap, cap - 1, 1 << 26, (1 << 26) + 1, 3, 1000003] + list(rng.integers(2, cap, 200)):
        n = int(n)
        k = np.arange(28, dtype=np.int64)
        c = (k[:, None] * n + np.arange(-3, 4)[None, :]).ravel()
        c = c[(c >= 0) & (c < 27 * n)]
        c = np.concatenate([c, rng.integers(0, 27 * n, 2000)])
        assert np.array_equal(canon(c, n), c % n)
