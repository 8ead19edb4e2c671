"""Named workloads: the synthetic Gaussian-mixture inputs and index shapes.

Every workload draws d = 128 float32 rows holding integers 0..255 from a
seeded mixture of 1000 Gaussians (centres uniform in [0, 128], sigma = 20 per
dimension, rounded and clipped); train, base and query rows come from separate
streams of the same generator.  Every index is trained with the k-means
iteration counts of the reference's ivfpq_new (50 coarse, 100 per
subquantizer; parallel/ivf_pq/ivf_training.cpp:77, parallel/pq-utils/pq_new.cpp:53).
"""
D = 128
NCOMP = 1000
SIGMA = 20.0
STREAM_TRAIN = 2
STREAM_BASE = 3
STREAM_QUERY = 4

WORKLOADS = {
    # reference CPU-path case: 10k base, 100 queries, 25k training vectors,
    # 256 lists, PQ m=8 x 8 bit, w=8, k=100; k-means iterations as ivfpq_new
    # (50 coarse, 100 per subquantizer)
    "ivfpq_10k_l256_m8": dict(seed=101, ntrain=25_000, nbase=10_000, nquery=100, nlist=256, m=8,
                              w=8, k=100, zipf=None, niter_coarse=50, niter_pq=100),
    # plain GPU drop-in: 1M base, 10k queries, 100k training, 1024 lists, m=8, w=16, k=100
    "ivfpq_1m_l1024_m8": dict(seed=102, ntrain=100_000, nbase=1_000_000, nquery=10_000, nlist=1024,
                              m=8, w=16, k=100, zipf=None, niter_coarse=50, niter_pq=100),
    # headline: 10M base (Zipf s=1.1 mixture weights), 4096 lists, m=16,
    # 10k queries, w=32, k=100; trained on 1M (= base/10)
    "ivfpq_10m_l4096_m16": dict(seed=103, ntrain=1_000_000, nbase=10_000_000, nquery=10_000,
                                nlist=4096, m=16, w=32, k=100, zipf=1.1, niter_coarse=50,
                                niter_pq=100),
    # serving loop: 100M base streamed in 1M-row chunks (ids = global row
    # numbers) and kept only as PQ16 codes; 16384 lists trained on 1M
    # vectors; open-loop Poisson requests, batch sizes log-uniform 1..1024,
    # w from {8, 32, 128}, k from {10, 100}
    "ivfpq_100m_l16384_m16": dict(seed=105, ntrain=1_000_000, nbase=100_000_000, nquery=10_000,
                                  nlist=16384, m=16, w=32, k=100, zipf=None, niter_coarse=50,
                                  niter_pq=100, stream_chunk=1_000_000, serve_w=(8, 32, 128),
                                  serve_k=(10, 100), serve_pool=100_000),
}
