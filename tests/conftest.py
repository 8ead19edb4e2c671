import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def mivf():
    """The product library; GPU tests fail loudly if it is missing."""
    import mestrado_amd as M
    M.lib()
    assert M.device_count() >= 1, "no HIP device visible"
    return M


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle as O
    O.lib()
    return O


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="session")
def c1(orc):
    """Small uniform-mixture workload (10k base, 100 queries, nlist 256, m 8)
    with the golden quantizers."""
    from mestrado_amd.workloads import D, NCOMP, SIGMA, STREAM_BASE, STREAM_QUERY, STREAM_TRAIN, WORKLOADS
    c = WORKLOADS["ivfpq_10k_l256_m8"]
    cent = orc.gen_centres(c["seed"], NCOMP, D)
    return dict(
        cfg=c, cent=cent,
        train=orc.gen_rows(c["seed"], STREAM_TRAIN, 0, c["ntrain"], cent, None, SIGMA),
        base=orc.gen_rows(c["seed"], STREAM_BASE, 0, c["nbase"], cent, None, SIGMA),
        query=orc.gen_rows(c["seed"], STREAM_QUERY, 0, c["nquery"], cent, None, SIGMA),
        gold_q=np.load(os.path.join(GOLD, "c1_quantizers.npz")),
    )


@pytest.fixture(scope="session")
def zipf16(mivf, orc):
    """Headline-shaped (Zipf s=1.1 mixture, m=16) index at reduced size:
    300k base, 300 queries, nlist 512.  Returns (index, oracle index tuple,
    queries, host base rows)."""
    from mestrado_amd.workloads import NCOMP, SIGMA, STREAM_BASE, STREAM_QUERY, STREAM_TRAIN
    seed = 103
    cent = orc.gen_centres(seed, NCOMP, 128)
    thr = orc.zipf_thresholds(seed, NCOMP, 1.1)
    train = orc.gen_rows(seed, STREAM_TRAIN, 0, 40_000, cent, thr, SIGMA)
    base = orc.gen_rows(seed, STREAM_BASE, 0, 300_000, cent, thr, SIGMA)
    query = orc.gen_rows(seed, STREAM_QUERY, 0, 300, cent, thr, SIGMA)
    coarse, pq = mivf.train(train, 512, 16, 256, 8, 8, seed)
    ix = mivf.Index(coarse, pq)
    ix.add(base)
    index = (coarse, pq) + tuple(ix.get_lists())
    return ix, index, query, base
