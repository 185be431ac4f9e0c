"""Morton order of particles (the library's own radix sort, sort.hip).

The keys are checked against a numpy restatement of the key function (bounding box,
one cubic scale 1023/extent, truncation to 10 bits per axis, bit interleave x, y,
z from the lowest bit), and the permutation against the definition of a stable
sort: a permutation of 0..n-1 whose keys never decrease and whose equal keys keep
ascending indices.  Sizes cover an empty call, one point, a workgroup tile edge
(4,096 keys per tile), ragged sizes and a million points on a coarse lattice, where
most keys repeat many times.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from omega3d_amd import _lib
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return _lib


def spread10(v):
    v = v.astype(np.uint32) & np.uint32(1023)
    v = (v | (v << np.uint32(16))) & np.uint32(0x030000FF)
    v = (v | (v << np.uint32(8))) & np.uint32(0x0300F00F)
    v = (v | (v << np.uint32(4))) & np.uint32(0x030C30C3)
    v = (v | (v << np.uint32(2))) & np.uint32(0x09249249)
    return v


def morton_keys(pos):
    pos = np.asarray(pos, np.float32)
    lo = pos.min(axis=1)
    ext = np.float32(max(pos[0].max() - lo[0], pos[1].max() - lo[1], pos[2].max() - lo[2]))
    sc = np.float32(1023.0) / ext if ext > 0 else np.float32(0.0)
    k = np.zeros(pos.shape[1], np.uint32)
    for d in range(3):
        q = ((pos[d] - lo[d]) * sc).astype(np.float32)
        c = np.clip(q, np.float32(0.0), np.float32(1023.0)).astype(np.uint32)
        k |= spread10(c) << np.uint32(d)
    return k


def run_order(L, pos):
    n = pos.shape[1]
    D = L.DeviceArray.from_host
    dpos = D(np.ascontiguousarray(pos, np.float32))
    dk = L.DeviceArray(max(n, 1), np.uint32)
    dp = L.DeviceArray(max(n, 1), np.int32)
    L.check(L.lib.o3d_morton_order_dev(C.c_void_p(dpos.row(0)), C.c_void_p(dpos.row(1)), C.c_void_p(dpos.row(2)),
                                        C.c_int64(n), C.c_void_p(dk.ptr), C.c_void_p(dp.ptr), None))
    L.synchronize()
    return dk.to_host()[:n], dp.to_host()[:n]


def check_stable(keys, perm):
    n = len(keys)
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.int32))
    ks = keys[perm]
    assert np.all(ks[1:] >= ks[:-1])
    ties = ks[1:] == ks[:-1]
    assert np.all(perm[1:][ties] > perm[:-1][ties])
    np.testing.assert_array_equal(perm, np.argsort(keys, kind="stable").astype(np.int32))


@pytest.mark.parametrize("n", [1, 2, 63, 4095, 4096, 4097, 70001])
def test_morton_order_random(L, n):
    rng = np.random.default_rng(1000 + n)
    pos = rng.normal(size=(3, n)).astype(np.float32)
    pos[2] *= 0.25                      # a flattened box: the scale comes from the widest axis
    keys, perm = run_order(L, pos)
    np.testing.assert_array_equal(keys, morton_keys(pos))
    check_stable(keys, perm)


def test_morton_order_many_ties(L):
    n = 1 << 20
    rng = np.random.default_rng(7)
    pos = (rng.integers(0, 24, size=(3, n)) / 23.0).astype(np.float32)   # 13,824 distinct keys
    keys, perm = run_order(L, pos)
    np.testing.assert_array_equal(keys, morton_keys(pos))
    check_stable(keys, perm)


def test_morton_order_degenerate(L):
    # every point equal: extent 0, every key 0, the identity order
    pos = np.ones((3, 5000), np.float32)
    keys, perm = run_order(L, pos)
    assert not keys.any()
    np.testing.assert_array_equal(perm, np.arange(5000, dtype=np.int32))
    # empty call
    k0, p0 = run_order(L, np.zeros((3, 0), np.float32))
    assert k0.size == 0 and p0.size == 0
