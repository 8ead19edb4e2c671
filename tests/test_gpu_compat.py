"""The reference-signature entry points, run on the GPU from C++ callers
compiled against include/mestrado_compat*.h (tests/compat/), compared bit for
bit with the oracle:

* parallel2/ (GPU fork): preallocate_gpu_mem -> core_gpu -> deallocate_gpu_mem
  as do_on() calls them (parallel2/ivf_pq/ivf_search.cpp:120-159), with lists
  shorter than k so that the min(n_q, k) packing at qid_to_starting_outid is
  exercised; ivfpq_new, write_cent / read_cent, ivfpq_assign and read_ivf.
* sequential/: ivfpq_search with RERANK (sequential/ivf_pq/ivf_search.cpp:9-153)
  through the fork's float *centroids[nsq] pqtipo.
* parallel/: per-list ivfpq_search + concatenation + my_k_min from several host
  threads, as parallel_search does (parallel/ivf_pq/ivf_search.cpp:97-137).
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def caller(name, mivf):
    mivf.lib()  # libmivf first, so the caller binds to the same loaded copy
    path = os.path.join(HERE, "compat", f"libcaller_{name}.so")
    assert os.path.exists(path), f"{path} missing: run __graft_entry__.build()"
    return ctypes.CDLL(path)


def P(a):
    return ctypes.c_void_p(a.ctypes.data)


def residuals(orc, coarse, query, w):
    ci, _ = orc.knn(coarse, query, w)
    res = (query[:, None, :] - coarse[ci]).reshape(-1, 128).astype(np.float32)
    return res, ci.reshape(-1).astype(np.int32)


def sparse_lists(rng, nlist, m, maxlen=40):
    """Lists of 0..maxlen vectors, a third of them empty, few distinct codes
    (exact ties)."""
    sizes = rng.integers(0, maxlen, nlist)
    sizes[::3] = 0
    off = np.zeros(nlist + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    n = int(off[-1])
    ids = rng.permutation(n * 3)[:n].astype(np.int32)
    codes = rng.integers(0, 6, (n, m)).astype(np.int32)
    return off, ids, codes


@pytest.mark.parametrize("lists", ["c1", "sparse"])
def test_core_gpu_do_on(mivf, orc, c1, lists):
    L = caller("parallel2", mivf)
    g = c1["gold_q"]
    coarse, pq = g["coarse"], g["pq"]
    nlist, m, ks = coarse.shape[0], pq.shape[0], pq.shape[1]
    if lists == "c1":
        ix = mivf.Index(coarse, pq)
        ix.add(c1["base"])
        off, ids, codes8 = ix.get_lists()
        codes = codes8.astype(np.int32)
        w, k = 8, 100
    else:
        off, ids, codes = sparse_lists(np.random.default_rng(31), nlist, m)
        w, k = 4, 100
    query = c1["query"]
    nq = query.shape[0]
    res, rid_to_ivf = residuals(orc, coarse, query, w)
    out_ids = np.zeros(nq * k, np.int32)
    out_dis = np.zeros(nq * k, np.float32)
    starts = np.zeros(nq, np.int32)
    sizes = np.zeros(nq, np.int32)
    total = L.run_core_gpu(nlist, m, ks, P(pq), P(off), P(ids), P(codes), nq, w, k, P(res), P(rid_to_ivf),
                           P(out_ids), P(out_dis), P(starts), P(sizes))
    index = (coarse, pq, off, ids, codes.astype(np.uint8))
    rids, rdis = orc.search_residual(index, res, rid_to_ivf, w, k)
    lens = np.diff(off)[rid_to_ivf.reshape(nq, w)].sum(1)
    assert np.array_equal(sizes, np.minimum(lens, k))
    assert total == sizes.sum()
    if lists == "sparse":
        assert (sizes < k).any() and (sizes == 0).any()
    for q in range(nq):
        s, n = starts[q], sizes[q]
        assert np.array_equal(out_ids[s:s + n], rids[q, :n]), q
        assert np.array_equal(out_dis[s:s + n].view(np.uint32), rdis[q, :n].view(np.uint32)), q


def test_build_calls_and_files(mivf, orc, c1, tmp_path):
    """ivfpq_new -> write_cent -> read_cent -> ivfpq_assign, then read_ivf of
    the per-list files: the same quantizers as the C ABI's training with
    ivfpq_new's iteration counts and seed (50, 100, seed 2), and the same lists
    as the C ABI's add."""
    L = caller("parallel2", mivf)
    L.run_build.restype = ctypes.c_int64
    L.run_read_ivf.restype = ctypes.c_int64
    train = np.ascontiguousarray(c1["train"][:6000])
    base = np.ascontiguousarray(c1["base"][:4000])
    coarsek, nsq = 32, 8
    coarse = np.zeros((coarsek, 128), np.float32)
    pq = np.zeros((nsq, 256, 16), np.float32)
    off = np.zeros(coarsek + 1, np.int64)
    ids = np.zeros(base.shape[0], np.int32)
    codes = np.zeros((base.shape[0], nsq), np.int32)
    fh, fc, fa = (str(tmp_path / n).encode() for n in ("file_ivfpq", "cent_ivfpq", "coa_ivfpq"))
    n = L.run_build(P(train), train.shape[0], P(base), base.shape[0], coarsek, nsq, fh, fc, fa, P(coarse), P(pq),
                    P(off), P(ids), P(codes))
    assert n == base.shape[0]
    rc, rp = mivf.train(train, coarsek, nsq, 256, 50, 100, 2)
    assert np.array_equal(coarse.view(np.uint32), rc.view(np.uint32))
    assert np.array_equal(pq.view(np.uint32), rp.view(np.uint32))
    ix = mivf.Index(rc, rp)
    ix.add(base)
    r_off, r_ids, r_codes = ix.get_lists()
    assert np.array_equal(off, r_off) and np.array_equal(ids, r_ids)
    assert np.array_equal(codes, r_codes.astype(np.int32))
    # the oracle's encode of the same rows
    a, c = orc.encode(rc, rp, base)
    lst = np.searchsorted(off, np.arange(n), side="right") - 1
    assert np.array_equal(a[ids], lst) and np.array_equal(c[ids].astype(np.int32), codes)
    # per-list files (genivf layout) read back through read_ivf
    d = tmp_path / "ivf"
    d.mkdir()
    M = mivf.lib()
    assert M.mivf_write_ivf_dir(str(d).encode(), coarsek, nsq, P(r_off), P(r_ids), P(r_codes)) == 0
    off2 = np.zeros_like(off)
    ids2 = np.zeros_like(ids)
    codes2 = np.zeros_like(codes)
    assert L.run_read_ivf(coarsek, nsq, str(d).encode(), P(off2), P(ids2), P(codes2)) == base.shape[0]
    assert np.array_equal(off2, off) and np.array_equal(ids2, ids) and np.array_equal(codes2, codes)


@pytest.mark.parametrize("kl,k", [(256, 100), (64, 10)])
def test_sequential_search_with_rerank(mivf, orc, c1, kl, k):
    L = caller("sequential", mivf)
    g = c1["gold_q"]
    coarse, pq = g["coarse"], g["pq"]
    ix = mivf.Index(coarse, pq)
    ix.add(c1["base"])
    off, ids, codes8 = ix.get_lists()
    codes = codes8.astype(np.int32)
    query, base = c1["query"], c1["base"]
    nq, w = query.shape[0], 8
    out_ids = np.zeros((nq, k), np.int32)
    out_dis = np.zeros((nq, k), np.float32)
    assert L.run_seq_search(coarse.shape[0], pq.shape[0], pq.shape[1], P(coarse), P(pq), P(off), P(ids), P(codes),
                            nq, P(query), base.shape[0], P(base), k, kl, w, P(out_ids), P(out_dis)) == 0
    rids, rdis = orc.search_rerank((coarse, pq, off, ids, codes8), base, query, w, kl, k)
    assert np.array_equal(out_ids, rids)
    assert np.array_equal(out_dis.view(np.uint32), rdis.view(np.uint32))


@pytest.mark.parametrize("lists,threads", [("c1", 4), ("sparse", 3)])
def test_parallel_fork_search_loop(mivf, orc, c1, lists, threads):
    L = caller("parallel", mivf)
    g = c1["gold_q"]
    coarse, pq = g["coarse"], g["pq"]
    nlist, m, ks = coarse.shape[0], pq.shape[0], pq.shape[1]
    if lists == "c1":
        ix = mivf.Index(coarse, pq)
        ix.add(c1["base"])
        off, ids, codes8 = ix.get_lists()
        codes = codes8.astype(np.int32)
        w, k = 8, 100
    else:
        off, ids, codes = sparse_lists(np.random.default_rng(32), nlist, m)
        w, k = 6, 100
    query = c1["query"][:40]
    nq = query.shape[0]
    res, coaidx = residuals(orc, coarse, query, w)
    out_ids = np.zeros((nq, k), np.int32)
    out_dis = np.zeros((nq, k), np.float32)
    out_n = np.zeros(nq, np.int32)
    gt = np.zeros(2, np.float64)
    assert L.run_par_search(nlist, m, ks, P(pq), P(off), P(ids), P(codes), nq, w, k, P(res), P(coaidx), threads,
                            P(out_ids), P(out_dis), P(out_n), P(gt)) == 0
    rids, rdis = orc.search_residual((coarse, pq, off, ids, codes.astype(np.uint8)), res, coaidx, w, k)
    lens = np.diff(off)[coaidx.reshape(nq, w)].sum(1)
    assert np.array_equal(out_n, np.minimum(lens, k))
    assert np.array_equal(out_ids, rids)
    assert np.array_equal(out_dis.view(np.uint32), rdis.view(np.uint32))
    assert gt[0] > 0.0


def test_parallel_fork_list_distances(mivf, orc, c1):
    """ivfpq_search alone: the distances of one list in list order equal the
    oracle's table + sumidxtab2."""
    L = caller("parallel", mivf)
    g = c1["gold_q"]
    coarse, pq = g["coarse"], g["pq"]
    ix = mivf.Index(coarse, pq)
    ix.add(c1["base"])
    off, ids, codes8 = ix.get_lists()
    codes = codes8.astype(np.int32)
    res, coaidx = residuals(orc, coarse, c1["query"][:3], 2)
    for j in range(res.shape[0]):
        lst = int(coaidx[j])
        n = int(off[lst + 1] - off[lst])
        oi = np.zeros(max(n, 1), np.int32)
        od = np.zeros(max(n, 1), np.float32)
        got = L.run_par_list(coarse.shape[0], pq.shape[0], pq.shape[1], P(pq), P(off), P(ids), P(codes), P(res[j]),
                             lst, P(oi), P(od))
        assert got == n
        ref = orc.adc(orc.tables(res[j], pq), codes8[off[lst]:off[lst + 1]])
        assert np.array_equal(oi[:n], ids[off[lst]:off[lst + 1]])
        assert np.array_equal(od[:n].view(np.uint32), ref.view(np.uint32))
        # the C ABI call underneath, directly
        assert np.array_equal(ix.adc(res[j], codes8[off[lst]:off[lst + 1]]).view(np.uint32), ref.view(np.uint32))


def test_my_k_min_is_exact(mivf):
    """my_k_min on inputs where the reference's heap goes wrong (sorted input)
    and on heavy ties: the k smallest by (distance, id)."""
    L = caller("parallel", mivf)
    rng = np.random.default_rng(33)
    cases = [np.arange(1, 8, dtype=np.float32), rng.integers(0, 5, 3000).astype(np.float32),
             rng.random(50000).astype(np.float32)]
    for dis in cases:
        n = dis.shape[0]
        ids = rng.permutation(n).astype(np.int32)
        for kt in sorted({1, min(3, n), min(100, n), min(1024, n)}):
            od = np.zeros(kt, np.float32)
            oi = np.zeros(kt, np.int32)
            assert L.run_kmin(n, P(dis), P(ids), kt, P(od), P(oi)) == 0
            order = np.lexsort((ids, dis))[:kt]
            assert np.array_equal(oi, ids[order]) and np.array_equal(od, dis[order])
    gi, gd = mivf.kmin(np.array([3.0, 1.0], np.float32), np.array([7, 9], np.int32), 4)
    assert gi.tolist() == [9, 7, -1, -1] and np.isinf(gd[2:]).all()
