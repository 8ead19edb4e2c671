// Reference-signature per-list calls of the MPI/OpenMP CPU fork (parallel/):
// ivfpq_search (parallel/ivf_pq/ivf_search.cpp:178-228) and my_k_min
// (parallel/ivf_pq/k_min.cpp:3) on top of the C ABI's mivf_adc / mivf_kmin.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/mestrado_compat_par.h"
#include "../../include/mivf.h"

namespace {

[[noreturn]] void die(const char *call) {
    std::fprintf(stderr, "Failed call: %s\nError: %s\n", call, mivf_last_error());
    std::exit(EXIT_FAILURE);
}
#define COMPAT(call)                       \
    do {                                   \
        if ((call) != MIVF_OK) die(#call); \
    } while (0)
#define COMPAT_REQUIRE(cond, msg)                                \
    do {                                                         \
        if (!(cond)) {                                           \
            std::fprintf(stderr, "Invalid argument: %s\n", msg); \
            std::exit(EXIT_FAILURE);                             \
        }                                                        \
    } while (0)

// grow-only device buffer
struct Buf {
    void *p = nullptr;
    size_t cap = 0;
    void *get(size_t want) {
        if (want > cap) {
            if (p) mivf_free(p);
            p = nullptr;
            cap = 0;
            COMPAT(mivf_malloc(&p, want));
            cap = want;
        }
        return p;
    }
    void release() {
        if (p) mivf_free(p);
        p = nullptr;
        cap = 0;
    }
};

struct DevList {
    void *codes = nullptr;  // [n][m] uint8
    int n = 0;
};

// Process-wide device state (the reference's search threads share one
// index); every call holds the mutex.
struct ParState {
    mivf_index *ix = nullptr;
    const float *pq_ptr = nullptr;
    int nsq = 0, ks = 0, ds = 0;
    std::unordered_map<const int *, DevList> lists;
    Buf res, dis, kdis, kids, koi, kod;
    void release() {
        for (auto &kv : lists) mivf_free(kv.second.codes);
        lists.clear();
        if (ix) mivf_index_free(ix);
        ix = nullptr;
        pq_ptr = nullptr;
        for (Buf *b : {&res, &dis, &kdis, &kids, &koi, &kod}) b->release();
    }
};
std::mutex g_mu;
ParState g;

void ensure_pq(const pqtipo &pq) {
    if (g.ix && g.pq_ptr == pq.centroids && g.nsq == pq.nsq && g.ks == pq.ks && g.ds == pq.ds) return;
    COMPAT_REQUIRE(pq.centroids && pq.nsq > 0 && pq.ds > 0 && pq.ks >= 1 && pq.ks <= 256,
                   "ivfpq_search: pqtipo needs centroids, nsq, ds > 0 and 1 <= ks <= 256 (8-bit codes)");
    g.release();
    const int d = pq.nsq * pq.ds;
    std::vector<float> zeros(d, 0.0f);  // the calls take residuals: no coarse quantizer needed
    COMPAT(mivf_index_create(d, 1, pq.nsq, pq.ks, zeros.data(), pq.centroids, &g.ix));
    g.pq_ptr = pq.centroids;
    g.nsq = pq.nsq;
    g.ks = pq.ks;
    g.ds = pq.ds;
}

const DevList &device_list(const ivf_t &L) {
    auto it = g.lists.find(L.codes.mat);
    if (it != g.lists.end() && it->second.n == L.codes.n) return it->second;
    const int n = L.codes.n, m = g.nsq;
    COMPAT_REQUIRE(L.codes.d == m, "ivfpq_search: ivf codes.d differs from pq.nsq");
    std::vector<uint8_t> c(size_t(n) * m);
    for (size_t i = 0; i < c.size(); ++i) {
        const int v = L.codes.mat[i];
        COMPAT_REQUIRE(v >= 0 && v < g.ks, "ivfpq_search: code out of range [0, ks)");
        c[i] = uint8_t(v);
    }
    DevList dl;
    dl.n = n;
    COMPAT(mivf_malloc(&dl.codes, c.size()));
    COMPAT(mivf_memcpy_htod(dl.codes, c.data(), c.size()));
    if (it != g.lists.end()) {
        mivf_free(it->second.codes);
        it->second = dl;
        return it->second;
    }
    return g.lists.emplace(L.codes.mat, dl).first->second;
}

double micros(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::micro>(b - a).count();
}

}  // namespace

dis_t ivfpq_search(ivf_t *ivf, float *residual, pqtipo pq, int centroid_idx, double *g1, double *g2) {
    COMPAT_REQUIRE(ivf && residual && centroid_idx >= 0, "ivfpq_search: null list array or residual");
    const ivf_t &L = ivf[centroid_idx];
    const int n = L.codes.n;
    COMPAT_REQUIRE(n >= 0 && L.idstam == n, "ivfpq_search: ivf idstam differs from codes.n");
    dis_t q;
    q.dis.n = n;
    q.dis.d = 1;
    q.dis.mat = static_cast<float *>(std::malloc(sizeof(float) * (n ? n : 1)));
    q.idx.n = n;
    q.idx.d = 1;
    q.idx.mat = static_cast<int *>(std::malloc(sizeof(int) * (n ? n : 1)));
    if (n == 0) return q;
    std::memcpy(q.idx.mat, L.ids, sizeof(int) * size_t(n));
    std::lock_guard<std::mutex> lk(g_mu);
    const auto t0 = std::chrono::steady_clock::now();
    ensure_pq(pq);
    const DevList &dl = device_list(L);
    const size_t rb = sizeof(float) * size_t(g.nsq) * g.ds;
    void *dres = g.res.get(rb);
    void *ddis = g.dis.get(sizeof(float) * size_t(n));
    COMPAT(mivf_memcpy_htod(dres, residual, rb));
    COMPAT(mivf_adc(g.ix, static_cast<const float *>(dres), static_cast<const uint8_t *>(dl.codes), n,
                    static_cast<float *>(ddis), nullptr));
    COMPAT(mivf_synchronize());
    const auto t1 = std::chrono::steady_clock::now();
    COMPAT(mivf_memcpy_dtoh(q.dis.mat, ddis, sizeof(float) * size_t(n)));
    const auto t2 = std::chrono::steady_clock::now();
    if (g1) *g1 += micros(t0, t1);
    if (g2) *g2 += micros(t1, t2);
    return q;
}

void my_k_min(dis_t q, int ktmp, float *dis, int *ids) {
    if (q.dis.n == 0 || ktmp == 0) return;
    COMPAT_REQUIRE(q.idx.n == q.dis.n, "my_k_min: q.idx.n differs from q.dis.n");
    COMPAT_REQUIRE(ktmp > 0 && ktmp <= q.dis.n && ktmp <= 1024, "my_k_min: ktmp must be in 1..min(n, 1024)");
    COMPAT_REQUIRE(dis && ids && q.dis.mat && q.idx.mat, "my_k_min: null pointer");
    std::lock_guard<std::mutex> lk(g_mu);
    const size_t n = size_t(q.dis.n);
    void *dd = g.kdis.get(sizeof(float) * n), *di = g.kids.get(sizeof(int) * n);
    void *oi = g.koi.get(sizeof(int) * ktmp), *od = g.kod.get(sizeof(float) * ktmp);
    COMPAT(mivf_memcpy_htod(dd, q.dis.mat, sizeof(float) * n));
    COMPAT(mivf_memcpy_htod(di, q.idx.mat, sizeof(int) * n));
    COMPAT(mivf_kmin(static_cast<const float *>(dd), static_cast<const int32_t *>(di), int64_t(n), ktmp,
                     static_cast<int32_t *>(oi), static_cast<float *>(od), nullptr));
    COMPAT(mivf_memcpy_dtoh(ids, oi, sizeof(int) * size_t(ktmp)));
    COMPAT(mivf_memcpy_dtoh(dis, od, sizeof(float) * size_t(ktmp)));
}

void mestrado_par_release(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g.release();
}
