"""GPU tests of the serving loop (continuous batching, mivf_server_*):
every request's ids and FP32 distances are bit-identical to the same
request searched alone, whatever the mix, slot count, batch limit, memory
budget or kernel family a micro-batch runs on (per-query or list-major,
with or without candidate overflow); the open-loop benchmark accounts for
every query."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_requests(query, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n):
        nq = int(rng.integers(1, 65))
        r0 = int(rng.integers(0, query.shape[0] - nq + 1))
        reqs.append((query[r0:r0 + nq], int(rng.choice([1, 8, 32, 64])), int(rng.choice([1, 10, 100]))))
    return reqs


@pytest.mark.parametrize("slots,max_batch,budget", [(3, 128, 0), (1, 64, 0), (4, 2048, 1)])
def test_serving_results_same_alone_and_mixed(mivf, zipf16, slots, max_batch, budget):
    ix, index, query, _ = zipf16
    reqs = make_requests(query, 40, seed=slots * 7 + max_batch)
    dq = [mivf.DeviceArray.from_numpy(q) for q, _, _ in reqs]
    sv = mivf.Server(ix, slots=slots, max_batch=max_batch, mem_budget=budget)
    tickets = [sv.submit(d, w, k) for d, (_, w, k) in zip(dq, reqs)]
    sv.drain()
    for t, (q, w, k) in zip(tickets, reqs):
        ids, dis = sv.result(t)
        aids, adis = ix.search(q, w, k)  # the same request alone
        assert np.array_equal(ids, aids) and np.array_equal(dis.view(np.uint32), adis.view(np.uint32)), \
            f"request nq={q.shape[0]} w={w} k={k} differs in the mix"
        assert sv.latency_ms(t) >= 0
    sv.close()


def test_serving_interleaved_submit_and_progress(mivf, zipf16):
    """Requests arriving while others are on the GPU are admitted as slots free up."""
    ix, index, query, _ = zipf16
    reqs = make_requests(query, 30, seed=5)
    sv = mivf.Server(ix, slots=2, max_batch=96)
    done, tickets = 0, []
    for d, (q, w, k) in zip([mivf.DeviceArray.from_numpy(q) for q, _, _ in reqs], reqs):
        tickets.append(sv.submit(d, w, k))
        done += sv.progress()
    sv.drain()
    for t, (q, w, k) in zip(tickets, reqs):
        ids, dis = sv.result(t)
        aids, adis = ix.search(q, w, k)
        assert np.array_equal(ids, aids) and np.array_equal(dis.view(np.uint32), adis.view(np.uint32))
    sv.close()


@pytest.mark.parametrize("cap", [16384, 64])
def test_serving_list_major_micro_batches(mivf, zipf16, cap):
    """Micro-batches of at least list_major_min_queries queries run on the
    list-major kernels; at candidate capacity 64 most of their queries
    overflow and are rerun by the per-query kernels after the batch.  Every
    request still matches itself searched alone on the per-query kernels."""
    ix, index, query, _ = zipf16
    reqs = make_requests(query, 40, seed=17 + cap)
    ix.set_option("list_major", 0)
    alone = [ix.search(q, w, k) for q, w, k in reqs]
    ix.set_option("list_major", 1)
    ix.set_option("list_major_min_queries", 48)
    ix.set_option("batched_capacity", cap)
    ix.set_option("batched_capacity_per_probe", 0 if cap == 64 else 256)
    try:
        dq = [mivf.DeviceArray.from_numpy(q) for q, _, _ in reqs]
        sv = mivf.Server(ix, slots=2, max_batch=256)
        sv.warm(query, 64, 100)
        assert sv.stats()["batches"] == 0
        tickets = [sv.submit(d, w, k) for d, (_, w, k) in zip(dq, reqs)]
        sv.drain()
        st = sv.stats()
        print("[serve] list-major micro-batches", st, flush=True)
        assert st["list_major_batches"] >= 1
        if cap == 64:
            assert st["rerun_queries"] >= 1
        for t, (q, w, k), (aids, adis) in zip(tickets, reqs, alone):
            ids, dis = sv.result(t)
            assert np.array_equal(ids, aids) and np.array_equal(dis.view(np.uint32), adis.view(np.uint32)), \
                f"request nq={q.shape[0]} w={w} k={k} differs in a list-major micro-batch"
        sv.close()
    finally:
        ix.set_option("list_major_min_queries", 2048)
        ix.set_option("batched_capacity", 16384)
        ix.set_option("batched_capacity_per_probe", 256)


def test_serving_min_fill_policy(mivf, zipf16):
    """With serve_min_fill set, a free slot waits for enough queries (or the
    oldest request's wait limit) while another micro-batch runs: fewer, larger
    micro-batches, the same rows."""
    ix, index, query, _ = zipf16
    reqs = make_requests(query, 40, seed=23)
    dq = [mivf.DeviceArray.from_numpy(q) for q, _, _ in reqs]
    ix.set_option("serve_min_fill", 200)
    ix.set_option("serve_max_wait_us", 2000)
    try:
        sv = mivf.Server(ix, slots=3, max_batch=512)
        tickets = [sv.submit(d, w, k) for d, (_, w, k) in zip(dq, reqs)]
        sv.drain()
        assert 1 <= sv.stats()["batches"] <= len(reqs)
        for t, (q, w, k) in zip(tickets, reqs):
            ids, dis = sv.result(t)
            aids, adis = ix.search(q, w, k)
            assert np.array_equal(ids, aids) and np.array_equal(dis.view(np.uint32), adis.view(np.uint32))
        sv.close()
    finally:
        ix.set_option("serve_min_fill", 0)
        ix.set_option("serve_max_wait_us", 0)


def test_serve_bench_accounts_for_every_query(mivf, zipf16):
    ix, index, query, _ = zipf16
    st = mivf.serve_bench(ix, query, rate=2000.0, nreq=200, seed=3, bmax=64, w_choices=(8, 32), k_choices=(10, 100))
    assert st["queries"] > 0 and st["p50_ms"] <= st["p99_ms"] and st["qps"] > 0 and st["batches"] >= 1
