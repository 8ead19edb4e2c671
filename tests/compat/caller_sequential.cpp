// A caller of the sequential fork's search with RERANK, written against
// include/mestrado_compat_seq.h: the quantizers and lists in the fork's own
// layouts (pqtipo with float *centroids[nsq], ivf_t with int32 codes), then
// ivfpq_search(ivfpq, ivf, vquery, vbase, k, kl, w, ids, dis) as
// sequential/ivf_pq/ivf_search.cpp:9 is called by sequential/ivf_test.cpp.
// The entry point is extern "C" only so that ctypes can call it.
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "mestrado_compat_seq.h"

extern "C" int run_seq_search(int nlist, int m, int ks, const float *coarse, const float *pqc, const int64_t *off,
                              const int32_t *ids, const int32_t *codes, int nq, const float *q, int nbase,
                              const float *base, int k, int kl, int w, int32_t *out_ids, float *out_dis) {
    const int d = 128, ds = d / m;
    ivfpq_t ivfpq;
    ivfpq.coarsek = ivfpq.coa_centroidsn = nlist;
    ivfpq.coa_centroidsd = d;
    ivfpq.coa_centroids = const_cast<float *>(coarse);
    ivfpq.pq.nsq = m;
    ivfpq.pq.ks = ks;
    ivfpq.pq.ds = ds;
    ivfpq.pq.centroidsn = ks;
    ivfpq.pq.centroidsd = ds;
    std::vector<float *> rows(m);
    for (int s = 0; s < m; ++s) rows[s] = const_cast<float *>(pqc) + size_t(s) * ks * ds;
    ivfpq.pq.centroids = rows.data();
    std::vector<ivf_t> ivf(nlist);
    for (int l = 0; l < nlist; ++l) {
        ivf[l].ids = const_cast<int *>(ids + off[l]);
        ivf[l].idstam = int(off[l + 1] - off[l]);
        ivf[l].codes.mat = const_cast<int *>(codes + off[l] * m);
        ivf[l].codes.n = ivf[l].idstam;
        ivf[l].codes.d = m;
    }
    mat vquery, vbase;
    vquery.mat = const_cast<float *>(q);
    vquery.n = nq;
    vquery.d = d;
    vbase.mat = const_cast<float *>(base);
    vbase.n = nbase;
    vbase.d = d;
    ivfpq_search(ivfpq, ivf.data(), vquery, vbase, k, kl, w, out_ids, out_dis);
    return 0;
}
