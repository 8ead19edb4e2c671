"""mivf_search_noalloc, called through the C ABI, against mivf_search: the
same ids and FP32 distances bit for bit, on two calls in a row (the second
with a smaller batch, so it reads what the first left in the shared scratch
array), on two index shapes; plus its two refusals (lists not yet on the
device, a batch larger than the fixed scratch)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def search_both(M, ix, q, w, k):
    L = M.lib()
    nq = q.shape[0]
    qd = M.DeviceArray.from_numpy(q)
    a_i, a_d = M.DeviceArray((nq, k), np.int32), M.DeviceArray((nq, k), np.float32)
    b_i, b_d = M.DeviceArray((nq, k), np.int32), M.DeviceArray((nq, k), np.float32)
    assert L.mivf_search_noalloc(ix.h, qd.ptr, nq, w, k, a_i.ptr, a_d.ptr, None) == 0, L.mivf_last_error()
    M.synchronize()
    assert L.mivf_search(ix.h, qd.ptr, nq, w, k, b_i.ptr, b_d.ptr, None) == 0, L.mivf_last_error()
    return a_i.numpy(), a_d.numpy(), b_i.numpy(), b_d.numpy()


def assert_same(a_i, a_d, b_i, b_d, what):
    assert np.array_equal(a_i, b_i), what
    assert np.array_equal(a_d.view(np.uint32), b_d.view(np.uint32)), what


def test_noalloc_equals_search_c1(mivf, c1):
    g = c1["gold_q"]
    ix = mivf.Index(g["coarse"], g["pq"])
    ix.add(c1["base"])
    q = c1["query"]
    L = mivf.lib()
    # lists added but not yet on the device: refused (uploading would allocate)
    qd = mivf.DeviceArray.from_numpy(q)
    oi, od = mivf.DeviceArray((q.shape[0], 10), np.int32), mivf.DeviceArray((q.shape[0], 10), np.float32)
    assert L.mivf_search_noalloc(ix.h, qd.ptr, q.shape[0], 8, 10, oi.ptr, od.ptr, None) == 1
    assert b"lists changed" in L.mivf_last_error()
    ix.search(q[:1], 1, 1)  # uploads the lists
    for nq, w, k in ((100, 8, 100), (37, 8, 100), (100, 64, 1000), (5, 1, 1)):
        r = search_both(mivf, ix, q[:nq], w, k)
        assert_same(*r, f"c1 nq={nq} w={w} k={k}")


def test_noalloc_equals_search_m16(mivf, zipf16):
    ix, _, query, _ = zipf16
    ix.search(query[:1], 1, 1)
    for nq, w, k in ((300, 32, 100), (120, 16, 10)):
        r = search_both(mivf, ix, query[:nq], w, k)
        assert_same(*r, f"m16 nq={nq} w={w} k={k}")


def test_noalloc_rejects_batches_above_the_scratch(mivf, c1):
    g = c1["gold_q"]
    ix = mivf.Index(g["coarse"], g["pq"])
    ix.add(c1["base"])
    ix.search(c1["query"][:1], 1, 1)
    q = mivf.DeviceArray.from_numpy(c1["query"][:1])
    oi, od = mivf.DeviceArray((1, 1), np.int32), mivf.DeviceArray((1, 1), np.float32)
    L = mivf.lib()
    # 200,000 queries at w = 256: far above 64 MiB; refused before any launch
    assert L.mivf_search_noalloc(ix.h, q.ptr, 200_000, 256, 1000, oi.ptr, od.ptr, None) == 3
    assert b"split the batch" in L.mivf_last_error()
