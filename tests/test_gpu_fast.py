"""GPU tests of the batched reduced-precision form (mivf_search_fast).

Its shortlist comes from 8-bit tables, so the shortlist itself is not the
reference's; what is pinned exactly is everything the caller sees: every
returned distance is the exact FP32 rerank distance of the returned id (the
oracle's distance arithmetic, bit for bit), rows are sorted by (distance, id),
ids are unique and come from the query's probed lists, results are
deterministic, and recall against FP64 brute force matches the exact RERANK
path's (sequential/ivf_pq/ivf_search.cpp:98-134) within a stated tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def check_exact_rows(orc, ids, dis, base, query, coarse_ids, list_of):
    nq, k = ids.shape
    for q in range(nq):
        row = ids[q]
        n = int((row >= 0).sum())
        assert (row[:n] >= 0).all() and (row[n:] == -1).all(), f"q={q}: padding not at the end"
        assert np.isinf(dis[q, n:]).all()
        assert len(set(row[:n].tolist())) == n, f"q={q}: duplicate ids"
        if n == 0:
            continue
        ex = orc.cross_dist(base[row[:n]], query[q:q + 1])[0]
        assert np.array_equal(ex.view(np.uint32), dis[q, :n].view(np.uint32)), \
            f"q={q}: distances are not the exact rerank distances"
        key = orc.ckey(dis[q, :n], row[:n])
        assert (np.diff(key.astype(np.uint64)) > 0).all(), f"q={q}: not sorted by (distance, id)"
        assert np.isin(list_of[row[:n]], coarse_ids[q]).all(), f"q={q}: id outside the probed lists"


def list_of_ids(off, ids, nbase):
    lo = np.full(nbase, -1, np.int64)
    for l in range(len(off) - 1):
        lo[ids[off[l]:off[l + 1]]] = l
    return lo


@pytest.fixture(scope="module")
def zipf16_gt(orc, zipf16):
    ix, index, query, base = zipf16
    return orc.bruteforce_f64(base, query, 1)[0][:, 0]


@pytest.mark.parametrize("w,kl,k,seed_lists", [(32, 256, 100, 1), (32, 128, 100, 2), (8, 1024, 10, 1),
                                                (64, 256, 1, 1), (1, 256, 100, 1), (18, 176, 100, 1),
                                                (17, 192, 100, 1)])
def test_fast_exact_rows_m16(mivf, orc, zipf16, zipf16_gt, w, kl, k, seed_lists):
    ix, index, query, base = zipf16
    gt = zipf16_gt
    ix.set_option("fast_seed_lists", seed_lists)
    ix.set_option("fast_capacity", 1 << 16)
    ix.set_option("profile", 1)
    try:
        ids, dis = ix.search_fast(query, w, kl, k, base)
        cands, over = ix.fast_stats()
        ids2, dis2 = ix.search_fast(query, w, kl, k, base)
    finally:
        ix.set_option("fast_seed_lists", 1)
        ix.set_option("fast_capacity", 1 << 16)  # the default
        ix.set_option("profile", 0)
    assert over == 0, f"{over} queries overflowed the candidate buffer"
    assert np.array_equal(ids, ids2) and np.array_equal(dis.view(np.uint32), dis2.view(np.uint32)), \
        "not deterministic"
    coarse_ids, _ = ix.coarse(query, w)
    lo = list_of_ids(index[2], index[3], base.shape[0])
    check_exact_rows(orc, ids, dis, base, query, coarse_ids, lo)
    # recall against FP64 brute force: within 1 query of the exact RERANK path
    rids, rdis = ix.search_rerank(query, w, kl, k, base)
    R = min(k, 100)
    rf = orc.recall_at(ids, gt[:, None], R)
    rr = orc.recall_at(rids, gt[:, None], R)
    # recall_at is in percent: one query is 100 / nq points
    assert rf >= rr - 100.0 / query.shape[0] - 1e-9, f"recall@{R}: fast {rf} vs rerank {rr}"
    # the two paths agree on nearly every returned id
    agree = np.mean([len(set(a[a >= 0]) & set(b[b >= 0])) / max(1, (b >= 0).sum()) for a, b in zip(ids, rids)])
    assert agree >= 0.97, f"only {agree:.3f} of the exact RERANK ids returned"


def test_fast_m8_and_empty_lists(mivf, orc, c1):
    """m = 8, ragged lists with empty ones (seed lists shorter than kl keep
    every vector), and k > available results (padding)."""
    q = c1["gold_q"]
    nlist = q["coarse"].shape[0]
    base = c1["base"]
    ix = mivf.Index(q["coarse"], q["pq"])
    ix.add(base)
    off, ids_all, codes = ix.get_lists()
    # drop every third list's members to make empty lists
    keep = np.ones(len(ids_all), bool)
    for l in range(0, nlist, 3):
        keep[off[l]:off[l + 1]] = False
    sizes = np.diff(off)
    sizes[::3] = 0
    off2 = np.zeros(nlist + 1, np.int64)
    off2[1:] = np.cumsum(sizes)
    ix2 = mivf.Index(q["coarse"], q["pq"])
    ix2.set_lists(off2, ids_all[keep], codes[keep])
    lo = list_of_ids(off2, ids_all[keep], base.shape[0])
    query = c1["query"]
    gt = orc.bruteforce_f64(base, query, 1)[0][:, 0]
    ix2.set_option("profile", 1)
    for w, kl, k in ((8, 256, 100), (2, 64, 64), (256, 512, 100), (1, 1024, 1000)):
        ids, dis = ix2.search_fast(query, w, kl, k, base)
        assert ix2.fast_stats()[1] == 0
        coarse_ids, _ = ix2.coarse(query, w)
        check_exact_rows(orc, ids, dis, base, query, coarse_ids, lo)
        # the shortlists differ from the exact RERANK path's only near the kl
        # boundary: with kl >= 2k nearly the same ids, and always the same
        # recall within one query
        rids, rdis = ix2.search_rerank(query, w, kl, k, base)
        agree = np.mean([len(set(a[a >= 0]) & set(b[b >= 0])) / max(1, (b >= 0).sum()) for a, b in zip(ids, rids)])
        assert kl < 2 * k or agree >= 0.97, f"w={w} kl={kl}: only {agree:.3f} of the exact RERANK ids returned"
        R = min(k, 100)
        rf, rr = orc.recall_at(ids, gt[:, None], R), orc.recall_at(rids, gt[:, None], R)
        assert rf >= rr - 100.0 / len(ids) - 1e-9, f"w={w} kl={kl}: recall@{R} fast {rf} vs rerank {rr}"


def test_fast_rejects_bad_shapes(mivf, c1):
    q = c1["gold_q"]
    ix = mivf.Index(q["coarse"], q["pq"])
    ix.add(c1["base"][:2000])
    with pytest.raises(mivf.MivfError):
        ix.search_fast(c1["query"], 8, 50, 100, c1["base"])   # k > kl
    with pytest.raises(mivf.MivfError):
        ix.search_fast(c1["query"], 8, 2048, 100, c1["base"])  # kl > 1024
    with pytest.raises(mivf.MivfError):
        ix.search_fast(c1["query"], 0, 64, 10, c1["base"])


def expected_tables(orc, coarse, pq, q, l):
    """Scaled reference tables x = T * 255 / bound (clipped to [0, 255]),
    bound = max_s (|r_s| + max_c |p_sc|)^2 * 1.0001, and their rounding."""
    m, ks, ds = pq.shape
    pmax = np.sqrt((pq.astype(np.float64) ** 2).sum(-1)).max(1) * (1 + 1e-6)
    x = np.zeros((m, ks, q.shape[0]), np.float64)
    bounds = []
    for g in range(q.shape[0]):
        r = (q[g] - coarse[l]).astype(np.float32)
        T = orc.tables(r, pq).astype(np.float64)  # [m][ks]
        a = np.sqrt((r.astype(np.float64).reshape(m, ds) ** 2).sum(1))
        bound = ((a + pmax) ** 2).max() * 1.0001
        bounds.append(bound)
        x[:, :, g] = np.clip(T * 255.0 / bound, 0, 255)
    return x, np.rint(x).astype(np.int64), np.array(bounds) / 255.0


@pytest.mark.parametrize("which", ["m16", "m8"])
def test_fast_tables_match_reference_arithmetic(mivf, orc, zipf16, c1, which):
    """The MFMA-built 8-bit tables: every entry within 0.75 quanta of the
    scaled reference table value (0.5 rounding + 0.25 FP16 operand rounding,
    the bound the candidate filter relies on), nearly all equal to its
    rounding, and the deltas agree."""
    if which == "m16":
        ix, index, query, _ = zipf16
        coarse, pq = index[0], index[1]
    else:
        q = c1["gold_q"]
        coarse, pq = q["coarse"], q["pq"]
        ix = mivf.Index(coarse, pq)
        query = c1["query"]
    probes, _ = ix.coarse(query[:16], 1)
    for n, l in ((16, int(probes[0, 0])), (5, int(probes[3, 0])), (16, 0)):
        t8, dl = ix.fast_tables(query[:n], l)
        x, et8, edl = expected_tables(orc, coarse, pq, query[:n], l)
        got = t8[:, :pq.shape[1], :n].astype(np.int64)
        err = np.abs(got - x)
        assert err.max() <= 0.75 + 1e-3, f"{which} list {l}: entry {err.max():.3f} quanta from the scaled table"
        same = (got == et8).mean()
        assert same >= 0.9, f"{which} list {l}: only {same:.4f} of the entries equal the rounded table"
        np.testing.assert_allclose(dl[:n], edl, rtol=1e-5)


def test_fast_attached_base(mivf, orc, c1):
    """A list-ordered copy of the raw vectors gives the same results as
    reranking against the caller's base; it is dropped when the lists change."""
    q = c1["gold_q"]
    base, query = c1["base"], c1["query"]
    ix = mivf.Index(q["coarse"], q["pq"])
    ix.add(base)
    with pytest.raises(mivf.MivfError):
        ix.search_fast(query, 8, 128, 100)          # nothing attached yet
    ix.attach_base(base)
    a_ids, a_dis = ix.search_fast(query, 8, 128, 100)
    b_ids, b_dis = ix.search_fast(query, 8, 128, 100, base)
    assert np.array_equal(a_ids, b_ids) and np.array_equal(a_dis.view(np.uint32), b_dis.view(np.uint32))
    off, ids, codes = ix.get_lists()
    ix.set_lists(off, ids, codes)
    with pytest.raises(mivf.MivfError):
        ix.search_fast(query, 8, 128, 100)          # stale copy dropped


@pytest.mark.parametrize("which,w,kl,k", [("c1", 8, 128, 100), ("c1", 256, 1024, 1000), ("m16", 32, 176, 100),
                                          ("m16", 17, 1024, 10)])
def test_fast_byte_base(mivf, orc, c1, zipf16, which, w, kl, k):
    """Option base_bytes: the attached copy of an integer-valued base holds one
    byte per value and the rerank returns the same ids and distances, bit for
    bit, as with the float copy, for integer-valued queries (integer dot
    products) and for queries that are not (FP32 sums); a base with a value
    that is not an integer in 0..255 is refused and leaves the index
    searchable."""
    if which == "c1":
        q = c1["gold_q"]
        base, query = c1["base"], c1["query"]
        ix = mivf.Index(q["coarse"], q["pq"])
        ix.add(base)
    else:
        ix, _, query, base = zipf16
    assert np.array_equal(base, np.clip(np.rint(base), 0, 255)), "the fixture base is integer-valued"
    ix.set_option("profile", 1)  # candidate statistics of every call
    qf = query.astype(np.float32).copy()
    qf[::2] += np.float32(0.375)       # every other query is not integer-valued
    qf[1::4, 5] = np.float32(256.0)    # and some are integers outside 0..255
    try:
        ix.attach_base(base)
        f_ids, f_dis = ix.search_fast(query, w, kl, k)
        assert ix.fast_stats()[1] == 0, "a query overflowed its candidate buffer (its shortlist is incomplete)"
        ff_ids, ff_dis = ix.search_fast(qf, w, kl, k)
        ix.set_option("base_bytes", 1)
        ix.attach_base(base)
        b_ids, b_dis = ix.search_fast(query, w, kl, k)
        assert ix.fast_stats()[1] == 0
        assert np.array_equal(f_ids, b_ids) and np.array_equal(f_dis.view(np.uint32), b_dis.view(np.uint32))
        bf_ids, bf_dis = ix.search_fast(qf, w, kl, k)
        assert np.array_equal(ff_ids, bf_ids) and np.array_equal(ff_dis.view(np.uint32), bf_dis.view(np.uint32))
        # the byte-row distances are the oracle's FP32 arithmetic, bit for bit
        for qs, ids_, dis_ in ((query, b_ids, b_dis), (qf, bf_ids, bf_dis)):
            for q in range(0, len(qs), 7):
                n = int((ids_[q] >= 0).sum())
                ex = orc.cross_dist(base[ids_[q, :n]], qs[q:q + 1])[0]
                assert np.array_equal(ex.view(np.uint32), dis_[q, :n].view(np.uint32)), f"q={q}"
        for bad_value in (0.5, -1.0, 256.0, np.nan):
            bad = base.copy()
            bad[base.shape[0] // 2, 7] = bad_value
            with pytest.raises(mivf.MivfError):
                ix.attach_base(bad)
            with pytest.raises(mivf.MivfError):
                ix.search_fast(query, w, kl, k)      # no copy after a refused one
        ix.set_option("base_bytes", 0)
        ix.attach_base(base)
        a_ids, a_dis = ix.search_fast(query, w, kl, k)
        assert np.array_equal(f_ids, a_ids) and np.array_equal(f_dis.view(np.uint32), a_dis.view(np.uint32))
    finally:
        ix.set_option("base_bytes", 0)
        ix.set_option("profile", 0)


@pytest.mark.parametrize("which,kl,seed_lists,cap", [("m16", 256, 1, 0), ("m16", 64, 2, 0), ("m8", 128, 1, 0),
                                                      ("m8", 1000, 1, 0), ("m16", 128, 1, 300), ("m16", 64, 2, 1024),
                                                      ("m8", 128, 1, 129), ("m16", 176, 1, 2048),
                                                      ("m8", 256, 2, 3000)])
def test_fast_seed_threshold(mivf, orc, zipf16, c1, which, kl, seed_lists, cap):
    """tau_q bounds the kl-th smallest reference ADC distance over the query's
    seed lists from above (the filter keeps everything up to it) and is
    tight (within 0.1%).  With a cap, the seed sees only the first cap
    vectors of its lists: tau then bounds the kl-th smallest of that sample
    (>= the kl-th of the whole seed lists) as tightly."""
    if which == "m16":
        ix, index, query, _ = zipf16
    else:
        q = c1["gold_q"]
        ix = mivf.Index(q["coarse"], q["pq"])
        ix.add(c1["base"])
        index = (q["coarse"], q["pq"]) + tuple(ix.get_lists())
        query = c1["query"]
    coarse, pq, off, ids, codes = index
    w = 16
    query = query[:40]
    ix.set_option("fast_seed_lists", seed_lists)
    ix.set_option("fast_seed_cap", cap)
    try:
        tau = ix.fast_seed_tau(query, w, kl)
    finally:
        ix.set_option("fast_seed_lists", 1)
        ix.set_option("fast_seed_cap", 2048)  # the library default
    probes, _ = ix.coarse(query, w)
    lens = np.diff(off)
    for qi in range(query.shape[0]):
        have, ns = 0, 0
        while ns < w and (ns < seed_lists or have < kl):
            have += lens[probes[qi, ns]]
            ns += 1
        vals = []
        for r in range(ns):
            l = probes[qi, r]
            if lens[l] == 0:
                continue
            tab = orc.tables((query[qi] - coarse[l]).astype(np.float32), pq)
            vals.append(orc.adc(tab, codes[off[l]:off[l + 1]]))
        allv = np.concatenate(vals) if vals else np.zeros(0, np.float32)
        sample = allv[:max(cap, kl)] if cap else allv
        vals = np.sort(sample)
        if len(vals) < kl:
            assert np.isinf(tau[qi]), f"q={qi}: fewer than kl seed vectors but finite tau"
            continue
        kth = float(vals[kl - 1])
        assert tau[qi] >= float(np.sort(allv)[kl - 1]), f"q={qi}: tau below the kl-th seed-list distance"
        assert tau[qi] >= kth, f"q={qi}: tau {tau[qi]} below the kl-th seed distance {kth}"
        assert tau[qi] <= kth * 1.001 + 1.0, f"q={qi}: tau {tau[qi]} loose against {kth}"


@pytest.mark.parametrize("case", ["distinct", "ties", "one_value", "few_values", "negative"])
def test_fast_select_picks_the_kl_smallest_keys(mivf, orc, case):
    """The shortlist step alone: the ids of the kl smallest (distance, position)
    keys of each query's candidates, whatever their spread.  Rows cover empty,
    shorter than kl, exactly kl, and more than the buffer (count clipped);
    heavy ties push the distance histogram past its LDS bin and into the
    all-keys radix select."""
    rng = np.random.default_rng(7)
    capq, kl = 6000, 100
    counts = [0, 1, kl - 1, kl, kl + 1, 700, 3000, capq, capq + 50, 4096]
    nq = len(counts)
    cand = np.zeros((nq, capq), np.uint64)
    for q, c in enumerate(counts):
        n = min(c, capq)
        if case == "distinct":
            d = rng.uniform(0.0, 1e5, n).astype(np.float32)
        elif case == "ties":
            d = rng.integers(0, 40, n).astype(np.float32) * np.float32(1.5)
        elif case == "one_value":
            d = np.full(n, 123.25, np.float32)
        elif case == "few_values":
            d = np.where(rng.random(n) < 0.02, rng.uniform(0, 10, n), 50.0).astype(np.float32)
        else:
            d = rng.normal(0.0, 1.0, n).astype(np.float32) * np.float32(1e3)
        pos = rng.permutation(1 << 20)[:n]
        cand[q, :n] = orc.ckey(d, pos)
        cand[q, n:] = rng.integers(0, 1 << 62, capq - n, dtype=np.uint64)  # beyond the count: ignored
    short = mivf.fast_select(cand, np.array(counts, np.int32), kl)
    for q, c in enumerate(counts):
        n = min(c, capq)
        want = np.sort(cand[q, :n])[:kl] & np.uint64(0xFFFFFFFF)
        got = short[q]
        m = min(n, kl)
        assert (got[m:] == -1).all(), f"{case} q={q}: padding"
        assert np.array_equal(np.sort(got[:m].astype(np.int64)), np.sort(want.astype(np.int64))), \
            f"{case} q={q} n={n}: not the kl smallest keys"


@pytest.mark.parametrize("front", [0, 1, 3, 31])
def test_fast_running_bound_changes_nothing(mivf, zipf16, front):
    """The running per-query bound (shared by the workgroups through
    candidate-distance histograms) only drops vectors that cannot enter the
    shortlist: results equal those without it, run after run, although the
    number of candidates varies with the schedule; so does taking the work
    items of each query's `front` nearest lists first."""
    ix, index, query, base = zipf16
    ix.set_option("fast_capacity", 1 << 16)
    ix.set_option("profile", 1)
    try:
        ix.set_option("fast_running_bound", 0)
        ref_ids, ref_dis = ix.search_fast(query, 32, 256, 100, base)
        ref_c = ix.fast_stats()[0]
        ix.set_option("fast_running_bound", 1)
        ix.set_option("fast_front_ranks", front)
        cands = []
        for _ in range(4):
            ids, dis = ix.search_fast(query, 32, 256, 100, base)
            cands.append(ix.fast_stats()[0])
            assert np.array_equal(ids, ref_ids) and np.array_equal(dis.view(np.uint32), ref_dis.view(np.uint32))
        # a wider batch after a narrower one: the two grouping ranges fill the work items anew
        n_ids, n_dis = ix.search_fast(query[:37], 7, 64, 10, base)
        ix.set_option("fast_running_bound", 0)
        r_ids, r_dis = ix.search_fast(query[:37], 7, 64, 10, base)
        assert np.array_equal(n_ids, r_ids) and np.array_equal(n_dis.view(np.uint32), r_dis.view(np.uint32))
    finally:
        ix.set_option("fast_running_bound", 1)
        ix.set_option("fast_front_ranks", 0)
        ix.set_option("fast_capacity", 1 << 16)  # the default
        ix.set_option("profile", 0)
    assert max(cands) <= ref_c


def test_fast_seed_slack(mivf, orc, zipf16, zipf16_gt):
    """The seed filter's slack (8-bit quanta above the seed bound that still
    pass): every setting is deterministic and returns exact rerank rows.
    With the running bound off, the candidates are exactly the vectors that
    pass the filter, so their count grows with the slack; 3m/4 + 2 = 14 keeps
    every vector within the seed bound, and the default (2) keeps its recall
    to within a few queries."""
    ix, index, query, base = zipf16
    gt = zipf16_gt
    coarse_ids, _ = ix.coarse(query, 32)
    lo = list_of_ids(index[2], index[3], base.shape[0])
    res = {}
    ix.set_option("profile", 1)
    try:
        ix.set_option("fast_running_bound", 0)
        for slack in (0, 2, 14):
            ix.set_option("fast_seed_slack", slack)
            ids, dis = ix.search_fast(query, 32, 256, 100, base)
            cands, over = ix.fast_stats()
            ids2, dis2 = ix.search_fast(query, 32, 256, 100, base)
            assert over == 0
            assert np.array_equal(ids, ids2) and np.array_equal(dis.view(np.uint32), dis2.view(np.uint32)), \
                f"slack {slack}: not deterministic"
            check_exact_rows(orc, ids, dis, base, query, coarse_ids, lo)
            res[slack] = (ids, cands)
    finally:
        ix.set_option("fast_seed_slack", 2)  # the default
        ix.set_option("fast_running_bound", 1)
        ix.set_option("profile", 0)
    assert res[0][1] <= res[2][1] <= res[14][1], [res[s][1] for s in (0, 2, 14)]
    r2 = orc.recall_at(res[2][0], gt[:, None], 100)
    r14 = orc.recall_at(res[14][0], gt[:, None], 100)
    assert r2 >= r14 - 3 * 100.0 / query.shape[0], f"recall@100 {r2} with slack 2 vs {r14} with 14"
    with pytest.raises(mivf.MivfError):
        ix.set_option("fast_seed_slack", 65)


def var_adc_bounds(orc, pq, res, codes, tables):
    """Exact ADC of codes for residual res (oracle tables + sumidxtab2) and
    the error bound of table type `tables` for each vector."""
    tab = orc.tables(res, pq)  # [m][ks]
    m, ks = tab.shape
    exact = orc.adc(tab, codes)
    ent = tab[np.arange(m)[None, :], codes.astype(np.int64)]  # [n][m] entries of each vector
    if tables == 1:
        # each FP16 entry within 2^-11 of itself; FP32 sums
        bound = 2.0 ** -11 * np.abs(ent).sum(1) * 1.01 + 1e-6 * np.abs(exact)
    else:
        rng = tab.max(1) - tab.min(1)  # per subquantizer over the codes < ks
        bound = np.full(len(codes), 0.5002 * float((rng / 255.0).sum())) + 1e-6 * (np.abs(exact) + tab.max(1).sum())
    return exact, bound


@pytest.mark.parametrize("tables", [1, 2])
@pytest.mark.parametrize("shape", ["m8", "m16"])
def test_var_tables_adc_within_bound(mivf, orc, c1, zipf16, tables, shape):
    """The FP16 and the per-subquantizer offset/scale 8-bit tables: every
    approximate ADC the scan loop computes is within its table type's proven
    error bound of the oracle's exact ADC (the bound the candidate filter
    uses), for 16 queries against several lists."""
    if shape == "m8":
        g = c1["gold_q"]
        coarse, pq = g["coarse"], g["pq"]
        ix = mivf.Index(coarse, pq)
        ix.add(c1["base"])
        off, ids, codes = ix.get_lists()
        query = c1["query"][:16]
    else:
        ix, index, query, _ = zipf16
        coarse, pq, off, ids, codes = index
        query = query[:16]
    lens = np.diff(off)
    lists = [int(l) for l in np.argsort(-lens)[:3]] + [int(np.nonzero(lens > 0)[0][0])]
    worst = 0.0
    for l in lists:
        n = int(lens[l])
        approx = ix.fast_var_adc(query, l, n, tables)
        cl = codes[off[l]:off[l + 1]]
        for gq in range(query.shape[0]):
            res = (query[gq] - coarse[l]).astype(np.float32)
            exact, bound = var_adc_bounds(orc, pq, res, cl, tables)
            err = np.abs(approx[:, gq].astype(np.float64) - exact.astype(np.float64))
            assert (err <= bound).all(), f"list {l} query {gq}: max err {err.max()} over bound {bound[err.argmax()]}"
            worst = max(worst, float((err / np.maximum(bound, 1e-30)).max()))
    assert worst > 0.0  # the entries are rounded: some error is there


@pytest.mark.parametrize("tables", [1, 2])
@pytest.mark.parametrize("w,kl,k", [(32, 256, 100), (8, 1024, 10), (17, 192, 100)])
def test_fast_var_tables_exact_rows(mivf, orc, zipf16, zipf16_gt, tables, w, kl, k):
    """End to end with the FP16 and the 8-bit per-subquantizer tables: exact
    rerank distances, (distance, id) order, ids from the probed lists, and
    recall within one query of the exact RERANK path."""
    ix, index, query, base = zipf16
    ix.set_option("fast_tables", tables)
    ix.set_option("fast_capacity", 1 << 16)
    ix.set_option("profile", 1)
    try:
        ids, dis = ix.search_fast(query, w, kl, k, base)
        over = ix.fast_stats()[1]
        ids2, dis2 = ix.search_fast(query, w, kl, k, base)
    finally:
        ix.set_option("fast_tables", 0)
        ix.set_option("fast_capacity", 1 << 16)  # the default
        ix.set_option("profile", 0)
    assert over == 0
    assert np.array_equal(ids, ids2) and np.array_equal(dis.view(np.uint32), dis2.view(np.uint32)), "not deterministic"
    coarse_ids, _ = ix.coarse(query, w)
    lo = list_of_ids(index[2], index[3], base.shape[0])
    check_exact_rows(orc, ids, dis, base, query, coarse_ids, lo)
    rids, _ = ix.search_rerank(query, w, kl, k, base)
    R = min(k, 100)
    rf = orc.recall_at(ids, zipf16_gt[:, None], R)
    rr = orc.recall_at(rids, zipf16_gt[:, None], R)
    assert rf >= rr - 100.0 / query.shape[0] - 1e-9, f"tables={tables}: recall@{R} {rf} vs rerank {rr}"
    agree = np.mean([len(set(a[a >= 0]) & set(b[b >= 0])) / max(1, (b >= 0).sum()) for a, b in zip(ids, rids)])
    assert kl < 2 * k or agree >= 0.97, f"tables={tables}: only {agree:.3f} of the exact RERANK ids returned"


@pytest.mark.parametrize("tables", [1, 2])
def test_fast_var_tables_m8_empty_lists(mivf, orc, c1, tables):
    """m = 8 with empty lists and k larger than the available results."""
    q = c1["gold_q"]
    nlist = q["coarse"].shape[0]
    base = c1["base"]
    ix = mivf.Index(q["coarse"], q["pq"])
    ix.add(base)
    off, ids_all, codes = ix.get_lists()
    keep = np.ones(len(ids_all), bool)
    for l in range(0, nlist, 3):
        keep[off[l]:off[l + 1]] = False
    sizes = np.diff(off)
    sizes[::3] = 0
    off2 = np.zeros(nlist + 1, np.int64)
    off2[1:] = np.cumsum(sizes)
    ix2 = mivf.Index(q["coarse"], q["pq"])
    ix2.set_lists(off2, ids_all[keep], codes[keep])
    ix2.set_option("fast_tables", tables)
    lo = list_of_ids(off2, ids_all[keep], base.shape[0])
    query = c1["query"]
    for w, kl, k in ((8, 256, 100), (1, 1024, 1000), (32, 1024, 100)):
        ids, dis = ix2.search_fast(query, w, kl, k, base)
        coarse_ids, _ = ix2.coarse(query, w)
        check_exact_rows(orc, ids, dis, base, query, coarse_ids, lo)
        rids, _ = ix2.search_rerank(query, w, kl, k, base)
        agree = np.mean([len(set(a[a >= 0]) & set(b[b >= 0])) / max(1, (b >= 0).sum()) for a, b in zip(ids, rids)])
        assert kl < 2 * k or agree >= 0.97, f"tables={tables} w={w} kl={kl}: agreement {agree:.3f}"
