"""GPU tests of the serving loop (continuous batching, mivf_server_*):
every request's ids and FP32 distances are bit-identical to the same
request searched alone, whatever the mix, slot count, batch limit or memory
budget; the open-loop benchmark accounts for every query."""
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


def test_serve_bench_accounts_for_every_query(mivf, zipf16):
    ix, index, query, _ = zipf16
    st = mivf.serve_bench(ix, query, rate=2000.0, nreq=200, seed=3, bmax=64, w_choices=(8, 32), k_choices=(10, 100))
    assert st["queries"] > 0 and st["p50_ms"] <= st["p99_ms"] and st["qps"] > 0 and st["batches"] >= 1
