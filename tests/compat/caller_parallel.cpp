// A caller of the MPI/OpenMP CPU fork's per-list calls (parallel/), written
// against include/mestrado_compat_par.h.  run_par_search is the inner loop of
// parallel_search (parallel/ivf_pq/ivf_search.cpp:97-137): for each query,
// ivfpq_search on each of its w probes, the results concatenated with
// realloc + memcpy, ktmp = min(total, k), then my_k_min.  Queries run on
// `threads` OpenMP-style host threads (std::thread here) as in the reference.
// The entry points are extern "C" only so that ctypes can call them.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "mestrado_compat_par.h"

namespace {

struct Fork {
    pqtipo pq;
    std::vector<ivf_t> ivf;
};

Fork make_fork(int nlist, int m, int ks, const float *pqc, const int64_t *off, const int32_t *ids,
               const int32_t *codes) {
    Fork f;
    f.pq.nsq = m;
    f.pq.ks = ks;
    f.pq.ds = 128 / m;
    f.pq.centroids = const_cast<float *>(pqc);
    f.pq.centroidsn = ks * (128 / m);
    f.pq.centroidsd = m;
    f.ivf.resize(nlist);
    for (int l = 0; l < nlist; ++l) {
        f.ivf[l].ids = const_cast<int *>(ids + off[l]);
        f.ivf[l].dis = nullptr;
        f.ivf[l].idstam = int(off[l + 1] - off[l]);
        f.ivf[l].codes.mat = const_cast<int *>(codes + off[l] * m);
        f.ivf[l].codes.n = f.ivf[l].idstam;
        f.ivf[l].codes.d = m;
    }
    return f;
}

}  // namespace

extern "C" {

// Rows of out_ids / out_dis are k wide; out_n[q] = ktmp of query q.
int run_par_search(int nlist, int m, int ks, const float *pqc, const int64_t *off, const int32_t *ids,
                   const int32_t *codes, int nq, int w, int k, const float *residual, const int32_t *coaidx,
                   int threads, int32_t *out_ids, float *out_dis, int32_t *out_n, double *g) {
    Fork f = make_fork(nlist, m, ks, pqc, off, ids, codes);
    std::vector<double> g1(threads, 0.0), g2(threads, 0.0);
    auto worker = [&](int my_omp_rank) {
        for (int i = my_omp_rank; i < nq; i += threads) {
            dis_t qt;
            qt.idx.mat = static_cast<int *>(std::malloc(sizeof(int)));
            qt.dis.mat = static_cast<float *>(std::malloc(sizeof(float)));
            qt.idx.n = qt.dis.n = 0;
            qt.idx.d = qt.dis.d = 1;
            for (int j = 0; j < w; j++) {
                dis_t q = ivfpq_search(f.ivf.data(), const_cast<float *>(residual) + size_t(i * w + j) * 128, f.pq,
                                       coaidx[i * w + j], &g1[my_omp_rank], &g2[my_omp_rank]);
                qt.idx.mat = static_cast<int *>(std::realloc(qt.idx.mat, sizeof(int) * (qt.idx.n + q.idx.n + 1)));
                qt.dis.mat =
                    static_cast<float *>(std::realloc(qt.dis.mat, sizeof(float) * (qt.dis.n + q.dis.n + 1)));
                std::memcpy(&qt.idx.mat[qt.idx.n], &q.idx.mat[0], sizeof(int) * q.idx.n);
                std::memcpy(&qt.dis.mat[qt.dis.n], &q.dis.mat[0], sizeof(float) * q.dis.n);
                qt.idx.n += q.idx.n;
                qt.dis.n += q.dis.n;
                std::free(q.dis.mat);
                std::free(q.idx.mat);
            }
            const int ktmp = qt.idx.n < k ? qt.idx.n : k;
            out_n[i] = ktmp;
            for (int c = 0; c < k; ++c) {
                out_ids[size_t(i) * k + c] = -1;
                out_dis[size_t(i) * k + c] = __builtin_huge_valf();
            }
            my_k_min(qt, ktmp, out_dis + size_t(i) * k, out_ids + size_t(i) * k);
            std::free(qt.dis.mat);
            std::free(qt.idx.mat);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < threads; ++t) th.emplace_back(worker, t);
    worker(0);
    for (auto &t : th) t.join();
    g[0] = g[1] = 0.0;
    for (int t = 0; t < threads; ++t) {
        g[0] += g1[t];
        g[1] += g2[t];
    }
    mestrado_par_release();
    return 0;
}

// One list's distances in list order (ivfpq_search alone); returns its length.
int run_par_list(int nlist, int m, int ks, const float *pqc, const int64_t *off, const int32_t *ids,
                 const int32_t *codes, const float *residual, int list, int32_t *out_ids, float *out_dis) {
    Fork f = make_fork(nlist, m, ks, pqc, off, ids, codes);
    double g1 = 0, g2 = 0;
    dis_t q = ivfpq_search(f.ivf.data(), const_cast<float *>(residual), f.pq, list, &g1, &g2);
    std::memcpy(out_ids, q.idx.mat, sizeof(int) * q.idx.n);
    std::memcpy(out_dis, q.dis.mat, sizeof(float) * q.dis.n);
    const int n = q.dis.n;
    std::free(q.dis.mat);
    std::free(q.idx.mat);
    mestrado_par_release();
    return n;
}

// my_k_min alone over given candidates.
int run_kmin(int n, const float *dis, const int32_t *ids, int ktmp, float *out_dis, int32_t *out_ids) {
    dis_t q;
    q.dis.mat = const_cast<float *>(dis);
    q.dis.n = n;
    q.dis.d = 1;
    q.idx.mat = const_cast<int *>(ids);
    q.idx.n = n;
    q.idx.d = 1;
    my_k_min(q, ktmp, out_dis, out_ids);
    return 0;
}

}  // extern "C"
