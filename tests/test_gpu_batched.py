"""GPU tests of the batched exact form (mivf_search_batched, mivf_plan_*):
per-query w and k, list-major persistent kernel, results bit-identical to the
plain per-query search (mivf_search_var), for every compiled work-item shape,
with the overflow fallback forced, and through a captured hipGraph replayed
several times with eager work between the replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def assert_same(ids, dis, rids, rdis, what):
    bad = np.nonzero((ids != rids).any(1) | (dis.view(np.uint32) != rdis.view(np.uint32)).any(1))[0]
    assert bad.size == 0, f"{what}: {bad.size} rows differ, first {bad[:5]}"


def uneven(nq, seed, wmax=64):
    rng = np.random.default_rng(seed)
    return rng.integers(1, wmax + 1, nq).astype(np.int32), rng.choice([1, 10, 100], nq).astype(np.int32)


@pytest.mark.parametrize("g,nt", [(8, 512), (8, 256), (4, 256), (4, 512)])
def test_batched_equals_plain_m16(mivf, zipf16, g, nt):
    ix, index, query, _ = zipf16
    w, k = uneven(query.shape[0], seed=g + nt)
    rids, rdis = ix.search_var(query, w, k, 100)
    ix.set_option("batched_group", g)
    ix.set_option("batched_threads", nt)
    try:
        ids, dis = ix.search_batched(query, w, k, 100)
    finally:
        ix.set_option("batched_group", 8)
        ix.set_option("batched_threads", 512)
    assert_same(ids, dis, rids, rdis, f"batched g={g} nt={nt}")


@pytest.mark.parametrize("order", [0, 1])
def test_item_order_changes_nothing(mivf, zipf16, order):
    """Numbering the list-major work items longest list first (option
    item_order) is a schedule only: the batched exact form, the list-major
    mivf_search and the reduced-precision form return what they return with
    the lists in index order, bit for bit."""
    ix, index, query, base = zipf16
    w, k = uneven(query.shape[0], seed=5)
    rids, rdis = ix.search_var(query, w, k, 100)
    ix.set_option("profile", 1)
    ix.set_option("item_order", 0)
    f0 = ix.search_fast(query, 24, 192, 100, base)
    ix.set_option("item_order", order)
    try:
        ids, dis = ix.search_batched(query, w, k, 100)
        assert_same(ids, dis, rids, rdis, f"batched, item order {order}")
        ix.set_option("list_major_min_queries", 1)
        ids, dis = ix.search(query, 32, 100)
        ix.set_option("list_major", 0)
        pids, pdis = ix.search(query, 32, 100)
        assert_same(ids, dis, pids, pdis, f"list-major search, item order {order}")
        ix.set_option("list_major", 1)
        f1 = ix.search_fast(query, 24, 192, 100, base)
        assert ix.fast_stats()[1] == 0
        assert_same(f1[0], f1[1], f0[0], f0[1], f"reduced precision, item order {order}")
    finally:
        ix.set_option("item_order", 0)
        ix.set_option("list_major", 1)
        ix.set_option("list_major_min_queries", 2048)
        ix.set_option("profile", 0)


@pytest.mark.parametrize("running", [0, 1, 2])
def test_exact_running_bound_changes_nothing(mivf, zipf16, running):
    """The list-major exact scan's running per-query bounds (candidate
    distance histograms shared by the workgroups) only drop vectors that
    cannot be among a query's k best: the batched form and the list-major
    mivf_search equal the per-query kernels bit for bit with or without them,
    for uneven and uniform batches, run after run."""
    ix, index, query, base = zipf16
    w, k = uneven(query.shape[0], seed=11)
    rids, rdis = ix.search_var(query, w, k, 100)
    ix.set_option("list_major", 0)
    pids, pdis = ix.search(query, 32, 100)
    ix.set_option("list_major", 1)
    ix.set_option("exact_running_bound", running)
    ix.set_option("list_major_min_queries", 1)
    try:
        for rep in range(3):
            ids, dis = ix.search_batched(query, w, k, 100)
            assert_same(ids, dis, rids, rdis, f"batched, running bound {running}, run {rep}")
            ids, dis = ix.search(query, 32, 100)
            assert_same(ids, dis, pids, pdis, f"list-major search, running bound {running}, run {rep}")
        # k = 1 and k = 1000 (the bound then needs one / a thousand kept vectors)
        for kk in (1, 1000):
            ix.set_option("list_major", 0)
            r1 = ix.search(query, 16, kk)
            ix.set_option("list_major", 1)
            l1 = ix.search(query, 16, kk)
            assert_same(l1[0], l1[1], r1[0], r1[1], f"list-major k={kk}, running bound {running}")
    finally:
        ix.set_option("exact_running_bound", 2)  # the default
        ix.set_option("list_major", 1)
        ix.set_option("list_major_min_queries", 2048)


def test_batched_overflow_fallback(mivf, zipf16):
    """A tiny candidate buffer overflows for most queries; the plain rerun of
    those queries keeps the results exact."""
    ix, index, query, _ = zipf16
    w, k = uneven(query.shape[0], seed=9)
    rids, rdis = ix.search_var(query, w, k, 100)
    ix.set_option("batched_capacity", 64)
    ix.set_option("batched_capacity_per_probe", 0)
    try:
        ids, dis = ix.search_batched(query, w, k, 100)
        over = ix.batched_stats()
    finally:
        ix.set_option("batched_capacity", 16384)  # the library defaults
        ix.set_option("batched_capacity_per_probe", 256)
    assert over > 0, "the tiny buffer should overflow"
    assert_same(ids, dis, rids, rdis, "batched with overflow")
    ids, dis = ix.search_batched(query, w, k, 100)
    assert ix.batched_stats() == 0
    assert_same(ids, dis, rids, rdis, "batched without overflow")


def test_batched_m8_ties_empty_lists(mivf, orc, c1):
    """m = 8 index with empty lists and few distinct codes (exact distance
    ties -> lower id first; fewer candidates than k -> padding)."""
    q = c1["gold_q"]
    nlist = q["coarse"].shape[0]
    rng = np.random.default_rng(11)
    sizes = rng.integers(0, 40, nlist)
    sizes[::3] = 0
    off = np.zeros(nlist + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    n = int(off[-1])
    ids = rng.permutation(n * 3)[:n].astype(np.int32)
    codes = rng.integers(0, 4, (n, 8)).astype(np.uint8)
    ix = mivf.Index(q["coarse"], q["pq"])
    ix.set_lists(off, ids, codes)
    query = c1["query"]
    for g, nt in ((8, 512), (16, 512), (4, 256)):
        ix.set_option("batched_group", g)
        ix.set_option("batched_threads", nt)
        w, k = uneven(query.shape[0], seed=g, wmax=64)
        k[::5] = 1000
        rids, rdis = ix.search_var(query, w, k, 1000)
        bids, bdis = ix.search_batched(query, w, k, 1000)
        assert_same(bids, bdis, rids, rdis, f"m8 ties g={g}")


def test_plan_graph_replays(mivf, zipf16):
    """Autotuned plan, captured graph, four replays with different queries, w
    and k, eager searches in between; every replay equals the eager result."""
    ix, index, query, _ = zipf16
    nq = 200
    plan = mivf.Plan(ix, nq, 64, 100, autotune=True)
    rng = np.random.default_rng(4)
    for rep in range(4):
        rows = rng.choice(query.shape[0], nq, replace=False)
        qs = np.ascontiguousarray(query[rows])
        w, k = uneven(nq, seed=100 + rep)
        ids, dis = plan.search(qs, w, k)
        rids, rdis = ix.search_var(qs, w, k, 100)  # eager work between replays
        assert_same(ids, dis, rids, rdis, f"graph replay {rep}")
    info = plan.info()
    assert info["group"] in (4, 8, 16) and info["threads"] in (256, 512) and info["tuned_ms"] > 0
    plan.close()


def test_exact_tables_and_scan_adc_bitexact(mivf, orc, zipf16):
    """The batched kernel's FP32 tables and the ADC distances of its scan loop
    equal the reference arithmetic (the oracle) bit for bit, per query slot."""
    ix, index, query, _ = zipf16
    coarse, pq, off, ids, codes = index
    sizes = np.diff(off)
    for l in (0, int(np.argmax(sizes)), int(np.nonzero(sizes)[0][-1])):
        t = ix.exact_tables(query[:8], l)
        d = ix.exact_adc(query[:8], l, int(sizes[l]))
        for g in range(8):
            ref = orc.tables((query[g] - coarse[l]).astype(np.float32), pq)
            assert np.array_equal(t[:, :, g].view(np.uint32), ref.view(np.uint32)), f"table list {l} slot {g}"
            adc = orc.adc(ref, codes[off[l]:off[l + 1]])
            assert np.array_equal(d[:, g].view(np.uint32), adc.view(np.uint32)), f"ADC list {l} slot {g}"


def test_plan_follows_list_changes(mivf, orc, c1):
    """A captured plan holds the lists' device addresses: after an add (which
    grows and moves the lists) and another search in between, the next
    plan.search re-captures and returns the results of the new lists."""
    q = c1["gold_q"]
    ix = mivf.Index(q["coarse"], q["pq"])
    base = c1["base"]
    ix.add(base[:3000])
    query = c1["query"]
    w, k = uneven(query.shape[0], seed=41, wmax=32)
    plan = mivf.Plan(ix, query.shape[0], 32, 100, autotune=False)
    try:
        a = plan.search(query, w, k)
        b = ix.search_var(query, w, k, 100)
        assert_same(a[0], a[1], b[0], b[1], "plan before the add")
        ix.add(base[3000:], id0=3000)
        ix.search(query[:5], 4, 10)  # uploads (and reallocates) the grown lists
        a = plan.search(query, w, k)
        index = (q["coarse"], q["pq"]) + tuple(ix.get_lists())
        r = orc.search_var(index, query, w, k, 100)
        assert_same(a[0], a[1], r[0], r[1], "plan after the add")
        a2 = plan.search(query, w, k)  # replay of the new graph
        assert_same(a2[0], a2[1], r[0], r[1], "plan replay after the add")
    finally:
        plan.close()
