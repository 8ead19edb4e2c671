"""Write the golden fixtures under tests/golden/ from the CPU oracle.

Run in the build container:  python tests/golden/make_golden.py
Outputs (all small):
  c1_quantizers.npz   coarse [256][128] and PQ [8][256][16] trained by the
                      oracle's k-means (50 coarse / 100 PQ iterations, as
                      ivfpq_new, parallel/ivf_pq/ivf_training.cpp:77 and
                      parallel/pq-utils/pq_new.cpp:111)
  c1_search.npz       plain search (w=8, k=100) ids/distances + coarse top-w
  c1_rerank.npz       rerank (w=8, kl=256, k=100) ids/distances
  refq_search.npz     search with the reference's trained 256-list quantizers
                      (ref/cent_ivfpq_256.bin, ref/coa_ivfpq_256.bin) over the
                      C1 base and the 100 siftsmall queries (ref/siftsmall_query.fvecs)
  manifest.json       seeds, shapes and SHA-256 digests of inputs and lists
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as O  # noqa: E402
from mestrado_amd.workloads import WORKLOADS, NCOMP, D, SIGMA, STREAM_TRAIN, STREAM_BASE, STREAM_QUERY  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_ref_quantizers():
    hdr = np.fromfile(os.path.join(HERE, "ref", "file_ivfpq_256.bin"), dtype=np.int32)
    nsq, ks, ds, cn, cd, coarsek, coan, coad = [int(v) for v in hdr]
    pq = np.fromfile(os.path.join(HERE, "ref", "cent_ivfpq_256.bin"), dtype=np.float32).reshape(nsq, ks, ds)
    coarse = np.fromfile(os.path.join(HERE, "ref", "coa_ivfpq_256.bin"), dtype=np.float32).reshape(coan, coad)
    return coarse, pq


def load_fvecs(path):
    raw = np.fromfile(path, dtype=np.int32)
    d = int(raw[0])
    return raw.reshape(-1, d + 1)[:, 1:].copy().view(np.float32)


def c1_data():
    c = WORKLOADS["ivfpq_10k_l256_m8"]
    cent = O.gen_centres(c["seed"], NCOMP, D)
    train = O.gen_rows(c["seed"], STREAM_TRAIN, 0, c["ntrain"], cent, None, SIGMA)
    base = O.gen_rows(c["seed"], STREAM_BASE, 0, c["nbase"], cent, None, SIGMA)
    query = O.gen_rows(c["seed"], STREAM_QUERY, 0, c["nquery"], cent, None, SIGMA)
    return c, cent, train, base, query


def main():
    c, cent, train, base, query = c1_data()
    coarse, pq = O.train(train, c["nlist"], c["m"], 256, c["niter_coarse"], c["niter_pq"], c["seed"])
    index = O.build_index(coarse, pq, base)
    ids, dis, ci, cd = O.search(index, query, c["w"], c["k"], want_coarse=True)
    rids, rdis = O.search_rerank(index, base, query, c["w"], 256, c["k"])
    np.savez_compressed(os.path.join(HERE, "c1_quantizers.npz"), coarse=coarse, pq=pq)
    np.savez_compressed(os.path.join(HERE, "c1_search.npz"), ids=ids, dis=dis, coaidx=ci, coadis=cd)
    np.savez_compressed(os.path.join(HERE, "c1_rerank.npz"), ids=rids, dis=rdis)

    rcoarse, rpq = load_ref_quantizers()
    rindex = O.build_index(rcoarse, rpq, base)
    sq = load_fvecs(os.path.join(HERE, "ref", "siftsmall_query.fvecs"))
    qids, qdis = O.search(rindex, sq, 8, 100)
    np.savez_compressed(os.path.join(HERE, "refq_search.npz"), ids=qids, dis=qdis)

    gt, _ = O.bruteforce_f64(base, query, 1)
    manifest = {
        "workload": "ivfpq_10k_l256_m8",
        "params": c,
        "sha256": {
            "centres": sha(cent), "train": sha(train), "base": sha(base), "query": sha(query),
            "c1_offsets": sha(index[2]), "c1_ids": sha(index[3]), "c1_codes": sha(index[4]),
            "refq_offsets": sha(rindex[2]), "refq_ids": sha(rindex[3]), "refq_codes": sha(rindex[4]),
        },
        "recall_at_100_plain": O.recall_at(ids, gt[:, 0], 100),
        "recall_at_100_rerank": O.recall_at(rids, gt[:, 0], 100),
    }
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    main()
