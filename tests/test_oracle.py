"""CPU pins of the oracle (no GPU).

Each oracle function is checked against something that does not retype its
formula: the reference's own sources compiled where they lie (oracle/_ref,
yael binheap/sorting + my_k_min + the recall statistics), closed forms
(integer data makes every distance exact), FP64 evaluations, brute force on
tiny inputs, invariants, and the golden fixtures.
"""
import ctypes
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from mestrado_amd.workloads import D, NCOMP, SIGMA, STREAM_BASE, STREAM_QUERY, STREAM_TRAIN, WORKLOADS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
REFLIB = os.path.join(ROOT, "oracle", "_ref", "libyaelref.so")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REFLIB):
        if not os.path.isdir("/root/reference"):
            pytest.skip("reference sources not present (oracle/_ref is built only where they are)")
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True, capture_output=True)
    return ctypes.CDLL(REFLIB)


# ---------------------------------------------------------------- keys
def test_keys_order(orc):
    rng = np.random.default_rng(0)
    d = np.concatenate([rng.standard_normal(2000).astype(np.float32) * 1e3,
                        np.array([0.0, -0.0, 1e-38, -1e-38, np.inf, -np.inf], np.float32)])
    ids = rng.integers(0, 1 << 30, d.size).astype(np.int32)
    keys = orc.ckey(d, ids)
    order = np.argsort(keys, kind="stable")
    ref = np.lexsort((ids, d))  # textbook (distance, id) order
    # -0 and +0 compare equal in float order, so both orders must agree on values
    assert np.array_equal(d[order], d[ref]) and np.array_equal(ids[order], ids[ref])
    assert orc.fkey(np.float32(-0.0)) == orc.fkey(np.float32(0.0))


# ---------------------------------------------------------------- generator
def test_generator_statistics_and_range(orc):
    cent = orc.gen_centres(5, NCOMP, D)
    assert cent.min() >= 0 and cent.max() < 128 and abs(cent.mean() - 64) < 1.0
    x = orc.gen_rows(5, 3, 0, 20000, cent, None, SIGMA)
    assert np.all(x == np.rint(x)) and x.min() >= 0 and x.max() <= 255
    # rows are deterministic and independent of how they are split
    y = np.concatenate([orc.gen_rows(5, 3, 0, 7000, cent, None, SIGMA),
                        orc.gen_rows(5, 3, 7000, 13000, cent, None, SIGMA)])
    assert np.array_equal(x, y)
    assert not np.array_equal(x[:100], orc.gen_rows(5, 4, 0, 100, cent, None, SIGMA))


def test_irwin_hall_normal():
    """The per-dimension deviate: sum of twelve 16-bit uniforms, centred and
    scaled, has mean 0 and variance (2^32-1)/2^32 (textbook Irwin-Hall)."""
    rng = np.random.default_rng(1)
    u = rng.integers(0, 1 << 16, (200000, 12)).sum(1)
    z = (u - 393210) / 65536.0
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.01


def test_zipf_thresholds(orc):
    t = orc.zipf_thresholds(3, NCOMP, 1.1)
    assert t[-1] == 1 << 53 and np.all(np.diff(t.astype(np.float64)) >= 0)
    w = np.diff(np.concatenate([[0], t.astype(np.float64)])) / 2.0 ** 53
    ws = np.sort(w)[::-1]
    ranks = np.arange(1, NCOMP + 1)
    expect = ranks ** -1.1 / np.sum(ranks ** -1.1)
    assert np.allclose(ws, expect, rtol=1e-9, atol=1e-15)


# ---------------------------------------------------------------- distances
def test_cross_dist_integer_exact(orc):
    """Integer vectors: norms and dot products are exact, so the formula must
    give the exact squared distance (a dropped term or wrong sign fails)."""
    rng = np.random.default_rng(2)
    a = rng.integers(0, 256, (40, 128)).astype(np.float32)
    b = rng.integers(0, 256, (30, 128)).astype(np.float32)
    out = orc.cross_dist(a, b)
    exact = ((b[:, None, :].astype(np.int64) - a[None, :, :].astype(np.int64)) ** 2).sum(-1)
    assert np.array_equal(out.astype(np.int64), exact)


def test_cross_dist_matches_fp64(orc):
    rng = np.random.default_rng(3)
    a = (rng.standard_normal((50, 16)) * 30).astype(np.float32)
    b = (rng.standard_normal((20, 16)) * 30).astype(np.float32)
    out = orc.cross_dist(a, b).astype(np.float64)
    exact = ((b[:, None, :].astype(np.float64) - a[None, :, :]) ** 2).sum(-1)
    scale = (a.astype(np.float64) ** 2).sum(1)[None, :] + (b.astype(np.float64) ** 2).sum(1)[:, None]
    assert np.all(np.abs(out - exact) <= 64 * np.finfo(np.float32).eps * scale)


def test_cross_dist_norm_roles(orc):
    """The a-side norm accumulates in float, the b-side in double
    (nn.c:129-145).  Inputs where the two accumulations differ tell a role
    swap apart."""
    a = np.array([[4097.0, 1.0, 1.0, 1.0]], np.float32)   # 4097^2 + 3 is not a float
    b = np.zeros((1, 4), np.float32)
    # a-side: float sum fl(fl(4097^2) + 1) ... ; dot = 0
    s = np.float32(0)
    for v in a[0]:
        s = np.float32(s + np.float32(v * v))
    got = orc.cross_dist(a, b)[0, 0]
    assert got == np.float32(s)                     # float accumulation
    assert got != np.float32(4097.0 ** 2 + 3)       # not the double one
    got2 = orc.cross_dist(b, a)[0, 0]               # roles swapped: b side is now a
    assert got2 == np.float32(4097.0 ** 2 + 3)


# ---------------------------------------------------------------- selection
def test_topk_textbook(orc):
    rng = np.random.default_rng(4)
    d = rng.integers(0, 50, 5000).astype(np.float32)   # many ties
    ids = rng.permutation(5000).astype(np.int32)
    keys = orc.ckey(d, ids)
    for k in (1, 7, 100, 5000, 6000):
        got = orc.topk_keys(keys, k)
        order = np.lexsort((ids, d))[:k]
        exp = keys[order]
        assert np.array_equal(got[:min(k, 5000)], exp)
        assert np.all(got[5000:] == np.iinfo(np.uint64).max)


def test_knn_vs_yael_binheap(orc, ref):
    """knn_full's selection is yael's fbinheap (binheap.c:65-172): on tie-free
    inputs the oracle's top-k equals fbinheap_addn_label_range + fbinheap_sort."""
    rng = np.random.default_rng(5)
    mat2 = rng.standard_normal((300, 16)).astype(np.float32) * 10
    mat1 = rng.standard_normal((40, 16)).astype(np.float32) * 10
    k = 25
    idx, dis = orc.knn(mat2, mat1, k)
    dists = orc.cross_dist(mat2, mat1)  # [40][300]
    ref.fbinheap_new.restype = ctypes.c_void_p
    ref.fbinheap_new.argtypes = [ctypes.c_int]
    ref.fbinheap_addn_label_range.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    ref.fbinheap_sort.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    ref.fbinheap_delete.argtypes = [ctypes.c_void_p]
    for i in range(mat1.shape[0]):
        row = np.ascontiguousarray(dists[i])
        assert np.unique(row).size == row.size
        h = ref.fbinheap_new(k)
        ref.fbinheap_addn_label_range(h, row.size, 0, row.ctypes.data)
        lab = np.empty(k, np.int32)
        val = np.empty(k, np.float32)
        ref.fbinheap_sort(h, lab.ctypes.data, val.ctypes.data)
        ref.fbinheap_delete(h)
        assert np.array_equal(lab, idx[i]) and np.array_equal(val, dis[i])


def test_fvec_k_min_agrees(orc, ref):
    """pq_search k_min -> yael fvec_k_min (sorting.c:180): same k smallest.
    n <= 3k takes the Hoare-select path (the one the sequential fork always
    uses).  The max-heap path (n > 3k) calls fbinheap_new, which sorting.c
    uses without a prototype, so on LP64 its pointer result is truncated and
    that path crashes when built from the sources as they are."""
    rng = np.random.default_rng(6)
    v = rng.standard_normal(140).astype(np.float32)
    k = 50
    perm = np.empty(k, np.int32)
    ref.fvec_k_min.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    ref.fvec_k_min(v.ctypes.data, v.size, perm.ctypes.data, k)
    got = orc.topk_keys(orc.ckey(v, np.arange(v.size)), k) & np.uint64(0xFFFFFFFF)
    assert np.array_equal(np.sort(perm), np.sort(got.astype(np.int32)))


def test_reference_my_k_min_defect(orc, ref):
    """Documents defect D1: the reference's my_k_min (parallel/ivf_pq/k_min.cpp)
    is not a top-k.  On the sorted input 1..7 it returns 1, 2, 6, ...; the
    oracle (and the GPU path) return 1, 2, 3."""
    class Mat(ctypes.Structure):
        _fields_ = [("mat", ctypes.c_void_p), ("n", ctypes.c_int), ("d", ctypes.c_int)]

    class DisT(ctypes.Structure):
        _fields_ = [("dis", Mat), ("idx", Mat)]

    vals = np.arange(1, 8, dtype=np.float32)
    ids = np.arange(7, dtype=np.int32)
    q = DisT(Mat(vals.ctypes.data, 7, 1), Mat(ids.ctypes.data, 7, 1))
    out_d = np.zeros(3, np.float32)
    out_i = np.zeros(3, np.int32)
    f = ref._Z8my_k_min5dis_tiPfPi
    f.argtypes = [DisT, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    f(q, 3, out_d.ctypes.data, out_i.ctypes.data)
    assert list(out_d) != [1.0, 2.0, 3.0]
    good = orc.topk_keys(orc.ckey(vals, ids), 3)
    assert [orc.lib() and float(x) for x in vals[(good & np.uint64(0xFFFFFFFF)).astype(np.int64)]] == [1, 2, 3]


def test_recall_matches_reference_stats(orc, ref, capfd):
    """R@R as pq_test_compute_stats (parallel/pq-utils/pq_test_compute_stats.cpp)."""
    rng = np.random.default_rng(7)
    nq, k = 200, 100
    ids = np.stack([rng.permutation(1000)[:k] for _ in range(nq)]).astype(np.int32)
    gt = rng.integers(0, 1000, nq).astype(np.int32)

    class MatI(ctypes.Structure):
        _fields_ = [("mat", ctypes.c_void_p), ("n", ctypes.c_int), ("d", ctypes.c_int)]

    f = ref._Z21pq_test_compute_statsPi4matIi
    f.argtypes = [ctypes.c_void_p, MatI, ctypes.c_int]
    capfd.readouterr()
    f(ids.ctypes.data, MatI(gt.ctypes.data, nq, 1), k)
    ctypes.CDLL(None).fflush(None)
    out = capfd.readouterr().out
    refvals = {int(l.split("=")[0].strip()[2:]): float(l.split("=")[1]) for l in out.strip().splitlines()}
    for R, v in refvals.items():
        assert abs(orc.recall_at(ids, gt, R) - v) < 1e-3, (R, v)


# ---------------------------------------------------------------- search pieces
def test_adc_is_sum_of_table_entries(orc):
    rng = np.random.default_rng(8)
    tab = (rng.standard_normal((8, 256)) * 100).astype(np.float32)
    codes = rng.integers(0, 256, (500, 8)).astype(np.uint8)
    out = orc.adc(tab, codes)
    for i in range(0, 500, 37):
        s = np.float32(0)
        for j in range(8):
            s = np.float32(s + tab[j, codes[i, j]])   # sumidxtab2 order, float accumulator
        assert out[i] == s
    terms = tab[np.arange(8), codes].astype(np.float64)
    assert np.all(np.abs(out - terms.sum(1)) <= 8 * np.finfo(np.float32).eps * np.abs(terms).sum(1))


def test_tables_are_subvector_distances(orc):
    rng = np.random.default_rng(9)
    pq = rng.integers(-20, 20, (8, 256, 16)).astype(np.float32)
    r = rng.integers(-60, 60, 128).astype(np.float32)
    tab = orc.tables(r, pq)
    exact = ((r.reshape(8, 1, 16).astype(np.int64) - pq.astype(np.int64)) ** 2).sum(-1)
    assert np.array_equal(tab.astype(np.int64), exact)


def small_index(orc, seed=10, nlist=32, m=8, nbase=3000):
    cent = orc.gen_centres(seed, 50, D)
    train = orc.gen_rows(seed, STREAM_TRAIN, 0, 4000, cent, None, SIGMA)
    base = orc.gen_rows(seed, STREAM_BASE, 0, nbase, cent, None, SIGMA)
    q = orc.gen_rows(seed, STREAM_QUERY, 0, 20, cent, None, SIGMA)
    coarse, pq = orc.train(train, nlist, m, 256, 5, 5, seed)
    return orc.build_index(coarse, pq, base), base, q


def test_search_phases_equals_search(orc):
    """The staged search (tables, ADC into a concatenated array, then the
    selection) returns exactly the fused search's results, and the four phase
    times are non-negative."""
    index, base, q = small_index(orc)
    for w, k in ((1, 5), (8, 100), (32, 700)):
        a_i, a_d = orc.search(index, q, w, k)
        b_i, b_d, ph = orc.search_phases(index, q, w, k, nthreads=2)
        assert np.array_equal(a_i, b_i) and np.array_equal(a_d.view(np.uint32), b_d.view(np.uint32))
        assert set(ph) == {"coarse", "tables", "adc", "select"} and min(ph.values()) >= 0.0


def test_search_all_lists_is_exhaustive_adc(orc):
    """w = nlist probes every list: the result is the exhaustive ADC top-k
    (brute force over every base vector with its own list's table)."""
    index, base, q = small_index(orc)
    coarse, pq, off, ids, codes = index
    k = 30
    got_i, got_d = orc.search(index, q, coarse.shape[0], k)
    for qi in range(q.shape[0]):
        allk = []
        for l in range(coarse.shape[0]):
            r = (q[qi] - coarse[l]).astype(np.float32)
            t = orc.tables(r, pq)
            dd = orc.adc(t, codes[off[l]:off[l + 1]])
            allk.append(orc.ckey(dd, ids[off[l]:off[l + 1]]))
        allk = np.sort(np.concatenate(allk))[:k]
        assert np.array_equal(got_i[qi], (allk & np.uint64(0xFFFFFFFF)).astype(np.int32))


def test_rerank_full_shortlist_is_bruteforce(orc):
    """With every list probed and a shortlist covering every vector, the
    rerank returns the exact nearest neighbours: on integer data the yael
    formula is exact, so ids and distances equal the FP64 brute force."""
    index, base, q = small_index(orc, nbase=900)
    k = 20
    ids, dis = orc.search_rerank(index, base, q, index[0].shape[0], 1024, k)
    bi, bd = orc.bruteforce_f64(base, q, k)
    assert np.array_equal(ids, bi) and np.array_equal(dis.astype(np.float64), bd)


def test_coarse_is_nearest_centroids(orc):
    index, base, q = small_index(orc)
    coarse = index[0]
    w = 8
    _, _, ci, cd = orc.search(index, q, w, 10, want_coarse=True)
    ex = ((q[:, None, :].astype(np.float64) - coarse[None]) ** 2).sum(-1)
    for i in range(q.shape[0]):
        top = np.argsort(ex[i], kind="stable")[:w]
        # the FP32 formula may swap near-ties only
        assert set(ci[i]) == set(top) or np.allclose(np.sort(ex[i])[w - 1], np.sort(ex[i])[w], rtol=1e-5)
        assert np.all(np.diff(cd[i]) >= 0)


def test_encode_is_argmin(orc):
    index, base, q = small_index(orc)
    coarse, pq = index[0], index[1]
    a, c = orc.encode(coarse, pq, base[:300])
    dc = ((base[:300, None, :].astype(np.float64) - coarse[None]) ** 2).sum(-1)
    assert np.all(dc[np.arange(300), a] <= dc.min(1) * (1 + 1e-5))
    r = (base[:300] - coarse[a]).astype(np.float32).reshape(300, 8, 16)
    dp = ((r[:, :, None, :].astype(np.float64) - pq[None]) ** 2).sum(-1)   # [300][8][256]
    chosen = np.take_along_axis(dp, c[:, :, None].astype(np.int64), 2)[..., 0]
    assert np.all(chosen <= dp.min(2) + 1e-3 * (1 + dp.min(2)))


def test_bucket_vs_ivec_sort_index(orc, ref):
    """List bucketing is the stable index sort of the assignment
    (ivec_sort_index, yael sorting.c:264-277)."""
    rng = np.random.default_rng(12)
    a = rng.integers(0, 17, 2000).astype(np.int32)
    c = rng.integers(0, 256, (2000, 4)).astype(np.uint8)
    off, ids, co = orc.bucket(17, a, c)
    perm = np.empty(2000, np.int32)
    ref.ivec_sort_index.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    ref.ivec_sort_index(a.ctypes.data, 2000, perm.ctypes.data)
    # yael's index sort is a qsort (not stable): compare per list as sets, and
    # ours must be ascending inside each list (ids ascending, sequential fork)
    for l in range(17):
        mine = ids[off[l]:off[l + 1]]
        assert np.all(np.diff(mine) > 0)
        assert set(mine) == set(perm[np.searchsorted(np.sort(a), l):np.searchsorted(np.sort(a), l, "right")])
    assert np.array_equal(co, c[ids])


def test_kmeans_fixed_point(orc):
    """After enough Lloyd steps the centroids are the means of their points."""
    rng = np.random.default_rng(13)
    x = np.concatenate([rng.normal(c, 1.0, (200, 4)) for c in (0, 20, 40, 60)]).astype(np.float32)
    cent = orc.kmeans(x, 4, 30, 7)
    a, _ = orc.knn(cent, x, 1)
    for j in range(4):
        assert np.allclose(cent[j], x[a[:, 0] == j].mean(0), atol=1e-4)


# ---------------------------------------------------------------- goldens
def test_golden_c1_reproduces(orc):
    """Regenerate the C1 workload and index and compare with the committed
    golden fixtures (quantizers, list digests, search and rerank results)."""
    man = json.load(open(os.path.join(GOLD, "manifest.json")))
    c = WORKLOADS["ivfpq_10k_l256_m8"]
    cent = orc.gen_centres(c["seed"], NCOMP, D)
    base = orc.gen_rows(c["seed"], STREAM_BASE, 0, c["nbase"], cent, None, SIGMA)
    query = orc.gen_rows(c["seed"], STREAM_QUERY, 0, c["nquery"], cent, None, SIGMA)
    assert sha(cent) == man["sha256"]["centres"]
    assert sha(base) == man["sha256"]["base"] and sha(query) == man["sha256"]["query"]
    g = np.load(os.path.join(GOLD, "c1_quantizers.npz"))
    index = orc.build_index(g["coarse"], g["pq"], base)
    assert sha(index[3]) == man["sha256"]["c1_ids"] and sha(index[4]) == man["sha256"]["c1_codes"]
    s = np.load(os.path.join(GOLD, "c1_search.npz"))
    ids, dis, ci, cd = orc.search(index, query, c["w"], c["k"], want_coarse=True)
    assert np.array_equal(ids, s["ids"]) and np.array_equal(dis.view(np.uint32), s["dis"].view(np.uint32))
    assert np.array_equal(ci, s["coaidx"])
    r = np.load(os.path.join(GOLD, "c1_rerank.npz"))
    rids, rdis = orc.search_rerank(index, base, query, c["w"], 256, c["k"])
    assert np.array_equal(rids, r["ids"]) and np.array_equal(rdis, r["dis"])
