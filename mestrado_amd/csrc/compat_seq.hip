// Reference-signature ivfpq_search of the sequential fork (with RERANK,
// sequential/ivf_pq/ivf_search.cpp:9-153) on top of the C ABI.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/mestrado_compat_seq.h"
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

struct DevMem {
    void *p = nullptr;
    explicit DevMem(size_t bytes) { COMPAT(mivf_malloc(&p, bytes)); }
    ~DevMem() { mivf_free(p); }
};

}  // namespace

void ivfpq_search(ivfpq_t ivfpq, ivf_t *ivf, mat vquery, mat vbase, int k, int kl, int w, int *ids, float *dis) {
    const int nl = ivfpq.coarsek, m = ivfpq.pq.nsq, ks = ivfpq.pq.ks, ds = ivfpq.pq.ds, d = m * ds;
    if (vquery.n <= 0) return;
    if (vquery.d != d || vbase.d != d) {
        std::fprintf(stderr, "ivfpq_search: query/base dimension differs from the quantizer\n");
        std::exit(EXIT_FAILURE);
    }
    std::vector<float> pq(size_t(ks) * d);
    for (int s = 0; s < m; ++s) std::memcpy(pq.data() + size_t(s) * ks * ds, ivfpq.pq.centroids[s], sizeof(float) * ks * ds);
    mivf_index *ix = nullptr;
    COMPAT(mivf_index_create(d, nl, m, ks, ivfpq.coa_centroids, pq.data(), &ix));
    std::vector<int64_t> off(nl + 1, 0);
    for (int l = 0; l < nl; ++l) off[l + 1] = off[l] + ivf[l].idstam;
    std::vector<int32_t> lid(off[nl] > 0 ? off[nl] : 1), codes(size_t(off[nl] > 0 ? off[nl] : 1) * m);
    for (int l = 0; l < nl; ++l) {
        if (!ivf[l].idstam) continue;
        std::memcpy(lid.data() + off[l], ivf[l].ids, sizeof(int) * ivf[l].idstam);
        std::memcpy(codes.data() + off[l] * m, ivf[l].codes.mat, sizeof(int) * size_t(ivf[l].idstam) * m);
    }
    COMPAT(mivf_index_set_lists_i32(ix, off.data(), lid.data(), codes.data()));
    const size_t qb = sizeof(float) * size_t(vquery.n) * d, bb = sizeof(float) * size_t(vbase.n) * d;
    {
        DevMem dq(qb), db(bb ? bb : 4), di(sizeof(int) * size_t(vquery.n) * k), dd(sizeof(float) * size_t(vquery.n) * k);
        COMPAT(mivf_memcpy_htod(dq.p, vquery.mat, qb));
        if (bb) COMPAT(mivf_memcpy_htod(db.p, vbase.mat, bb));
        COMPAT(mivf_search_rerank(ix, static_cast<float *>(dq.p), vquery.n, w, kl, k, static_cast<float *>(db.p),
                                  vbase.n, static_cast<int32_t *>(di.p), static_cast<float *>(dd.p), nullptr));
        COMPAT(mivf_memcpy_dtoh(ids, di.p, sizeof(int) * size_t(vquery.n) * k));
        if (dis) COMPAT(mivf_memcpy_dtoh(dis, dd.p, sizeof(float) * size_t(vquery.n) * k));
    }
    mivf_index_free(ix);
}
