"""GPU parity at the full sizes of the two bit-exact workloads.

* ivfpq_1m_l1024_m8: 1,000,000 base, 100,000 training and 10,000 query
  vectors from the uniform mixture, IVF with 1,024 lists, PQ m=8, w=16, k=100.
* ivfpq_10m_l4096_m16: 10,000,000 base vectors (Zipf s=1.1 mixture weights),
  1,000,000 training vectors, IVF with 4,096 lists, PQ m=16, 10,000 queries:
  the uniform w=32, k=100 batch and the uneven batch (w ~ U{1..64},
  k ~ {1, 10, 100}) through the plain kernels, the batched list-major kernel
  and its graph-captured plan.

Each index is trained and encoded on the GPU exactly as the benchmark builds
it; the oracle then searches the same lists on the host cores and every row
must match bit for bit in ids and FP32 distances (the order and arithmetic of
parallel/ivf_pq/ivf_search.cpp:97-137 with ties broken by lower id as in
parallel2/ivf_pq/topk.cu:26-32).  A sample of the base rows is also encoded by
the oracle and must land in the same list with the same codes.
"""
import time

import numpy as np
import pytest

from mestrado_amd.workloads import D, NCOMP, SIGMA, STREAM_BASE, STREAM_QUERY, STREAM_TRAIN, WORKLOADS

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[fullshape]", *a, flush=True)


def assert_same_results(ids, dis, rids, rdis, what=""):
    assert ids.shape == rids.shape, what
    bad = np.nonzero((ids != rids).any(1) | (dis.view(np.uint32) != rdis.view(np.uint32)).any(1))[0]
    assert bad.size == 0, f"{what}: {bad.size} rows differ, first {bad[:5]}; " \
                          f"gpu {ids[bad[0], :5]} {dis[bad[0], :5]} vs oracle {rids[bad[0], :5]} {rdis[bad[0], :5]}"


def build_workload(M, name):
    """Train and fill the index of workload `name` on the device (the
    benchmark's recipe).  Returns (Index, oracle index tuple, host queries,
    device base rows)."""
    c = WORKLOADS[name]
    t0 = time.time()
    cent = M.synth_centres(c["seed"], NCOMP, D)
    thr = M.zipf_thresholds(c["seed"], NCOMP, c["zipf"]) if c["zipf"] else None
    train = M.synth_rows(c["seed"], STREAM_TRAIN, 0, c["ntrain"], cent, thr, SIGMA)
    coarse, pq = M.train(train, c["nlist"], c["m"], 256, c["niter_coarse"], c["niter_pq"], c["seed"])
    train.free()
    _say(f"{name}: trained in {time.time() - t0:.1f}s")
    base = M.synth_rows(c["seed"], STREAM_BASE, 0, c["nbase"], cent, thr, SIGMA)
    query = M.synth_rows(c["seed"], STREAM_QUERY, 0, c["nquery"], cent, thr, SIGMA).numpy()
    ix = M.Index(coarse, pq)
    ix.add(base)
    index = (coarse, pq) + tuple(ix.get_lists())
    _say(f"{name}: {ix.ntotal} vectors encoded, {time.time() - t0:.1f}s")
    return ix, index, query, base


def check_encode_sample(orc, index, base, n):
    """The first n base rows, encoded by the oracle, sit in the lists the GPU
    built, with the same codes (ids ascend inside a list)."""
    coarse, pq, off, ids, codes = index
    x = base.head(n)
    ra, rc = orc.encode(coarse, pq, x)
    pos = np.full(n, -1, np.int64)
    sel = ids < n
    pos[ids[sel]] = np.nonzero(sel)[0]
    assert (pos >= 0).all(), "sampled ids missing from the lists"
    lst = np.searchsorted(off, pos, side="right") - 1
    assert np.array_equal(lst, ra), "coarse assignment differs from the oracle"
    assert np.array_equal(codes[pos], rc), "PQ codes differ from the oracle"


@pytest.fixture(scope="module")
def c2(mivf):
    ix, index, query, base = build_workload(mivf, "ivfpq_1m_l1024_m8")
    yield ix, index, query, base
    ix.close()
    base.free()


@pytest.fixture(scope="module")
def c3(mivf):
    ix, index, query, base = build_workload(mivf, "ivfpq_10m_l4096_m16")
    yield ix, index, query, base
    ix.close()
    base.free()


def test_c2_encode_sample(orc, c2):
    _, index, _, base = c2
    check_encode_sample(orc, index, base, 20_000)


def test_c2_plain_full(orc, c2):
    """1M / IVF1024 / PQ8, all 10,000 queries, w=16, k=100."""
    ix, index, query, _ = c2
    c = WORKLOADS["ivfpq_1m_l1024_m8"]
    ids, dis = ix.search(query, c["w"], c["k"])
    t0 = time.time()
    rids, rdis = orc.search(index, query, c["w"], c["k"])
    _say(f"c2 oracle search {time.time() - t0:.1f}s")
    assert_same_results(ids, dis, rids, rdis, "c2 w=16 k=100 (list-major)")
    assert (ids >= 0).all()
    ix.set_option("list_major", 0)
    try:
        pids, pdis = ix.search(query, c["w"], c["k"])
    finally:
        ix.set_option("list_major", 1)
    assert_same_results(pids, pdis, rids, rdis, "c2 w=16 k=100 (per-query kernels)")


def test_c3_encode_sample(orc, c3):
    _, index, _, base = c3
    check_encode_sample(orc, index, base, 8_000)


def test_c3_headline_full(orc, c3):
    """10M / IVF4096 / PQ16, all 10,000 queries, w=32, k=100: the plain
    kernels and the batched list-major kernels against the oracle."""
    ix, index, query, _ = c3
    c = WORKLOADS["ivfpq_10m_l4096_m16"]
    w, k = c["w"], c["k"]
    ids, dis = ix.search(query, w, k)
    t0 = time.time()
    rids, rdis = orc.search(index, query, w, k)
    _say(f"c3 oracle search {time.time() - t0:.1f}s")
    assert_same_results(ids, dis, rids, rdis, "c3 plain w=32 k=100 (list-major)")
    ix.set_option("list_major", 0)
    try:
        pids, pdis = ix.search(query, w, k)
    finally:
        ix.set_option("list_major", 1)
    assert_same_results(pids, pdis, rids, rdis, "c3 plain w=32 k=100 (per-query kernels)")
    nq = query.shape[0]
    wq = np.full(nq, w, np.int32)
    kq = np.full(nq, k, np.int32)
    bids, bdis = ix.search_batched(query, wq, kq, k)
    assert_same_results(bids, bdis, rids, rdis, "c3 batched w=32 k=100")
    assert ix.batched_stats() == 0


def test_c3_uneven_full(mivf, orc, c3):
    """The uneven batch (w ~ U{1..64}, k ~ {1, 10, 100}): per-query plain
    kernels, the batched kernel and the autotuned graph plan (two replays)."""
    ix, index, query, _ = c3
    nq = query.shape[0]
    rng = np.random.default_rng(WORKLOADS["ivfpq_10m_l4096_m16"]["seed"])
    wq = rng.integers(1, 65, nq).astype(np.int32)
    kq = rng.choice([1, 10, 100], nq).astype(np.int32)
    t0 = time.time()
    rids, rdis = orc.search_var(index, query, wq, kq, 100)
    _say(f"c3 oracle uneven search {time.time() - t0:.1f}s")
    ids, dis = ix.search_var(query, wq, kq, 100)
    assert_same_results(ids, dis, rids, rdis, "c3 search_var")
    bids, bdis = ix.search_batched(query, wq, kq, 100)
    assert_same_results(bids, bdis, rids, rdis, "c3 search_batched")
    plan = mivf.Plan(ix, nq, 64, 100, autotune=True)
    try:
        for rep in range(2):
            pids, pdis = plan.search(query, wq, kq)
            assert_same_results(pids, pdis, rids, rdis, f"c3 plan replay {rep}")
    finally:
        plan.close()
