// A caller of the GPU fork's entry points (parallel2/), written against
// include/mestrado_compat.h as the reference's own code calls them:
//   run_core_gpu  -- do_on() of parallel2/ivf_pq/ivf_search.cpp:120-159:
//                    per-query output sizes min(n_q, k) and start offsets,
//                    then preallocate_gpu_mem -> core_gpu -> deallocate_gpu_mem
//   run_build     -- ivfpq_new, write_cent / read_cent, ivfpq_assign
//                    (parallel2/ivf_pq/ivf_training.cpp, train.cpp, genivf.cpp)
//   run_read_ivf  -- read_ivf of per-list files (parallel2/ivf_pq/ivf_search.cpp:293)
// The entry points are extern "C" only so that ctypes can call them.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mestrado_compat.h"

namespace {

std::vector<ivf_t> make_lists(int nlist, int m, const int64_t *off, const int32_t *ids, const int32_t *codes) {
    std::vector<ivf_t> ivf(nlist);
    for (int l = 0; l < nlist; ++l) {
        ivf[l].ids = const_cast<int *>(ids + off[l]);
        ivf[l].idstam = int(off[l + 1] - off[l]);
        ivf[l].codes.mat = const_cast<int *>(codes + off[l] * m);
        ivf[l].codes.n = ivf[l].idstam;
        ivf[l].codes.d = m;
    }
    return ivf;
}

int64_t flatten(const ivf_t *ivf, int nlist, int m, int64_t *off, int32_t *ids, int32_t *codes) {
    off[0] = 0;
    for (int l = 0; l < nlist; ++l) {
        const int n = ivf[l].idstam;
        std::memcpy(ids + off[l], ivf[l].ids, sizeof(int) * size_t(n));
        std::memcpy(codes + off[l] * m, ivf[l].codes.mat, sizeof(int) * size_t(n) * m);
        off[l + 1] = off[l] + n;
    }
    return off[nlist];
}

}  // namespace

extern "C" {

// Returns the number of packed results; starts[q] / sizes[q] locate query q's
// results in out_ids / out_dis (capacity nq * k).
int run_core_gpu(int nlist, int m, int ks, const float *pqc, const int64_t *off, const int32_t *ids,
                 const int32_t *codes, int nq, int w, int k, const float *residual, const int32_t *rid_to_ivf,
                 int32_t *out_ids, float *out_dis, int32_t *starts, int32_t *sizes) {
    const int d = 128;
    pqtipo PQ;
    PQ.nsq = m;
    PQ.ks = ks;
    PQ.ds = d / m;
    PQ.centroids = const_cast<float *>(pqc);
    PQ.centroidsn = ks * (d / m);
    PQ.centroidsd = m;
    std::vector<ivf_t> ivf = make_lists(nlist, m, off, ids, codes);
    // do_on: sizes and starting offsets of every query's results
    std::vector<int> qid_to_starting_outid(nq);
    int rid = 0, outid = 0;
    for (int qid = 0; qid < nq; ++qid) {
        int numImgs = 0;
        for (int i = 0; i < w; ++i, ++rid) numImgs += ivf[rid_to_ivf[rid]].idstam;
        const int size = numImgs < k ? numImgs : k;
        sizes[qid] = size;
        starts[qid] = outid;
        qid_to_starting_outid[qid] = outid;
        outid += size;
    }
    matI idxs;
    mat dists;
    idxs.mat = new int[outid > 0 ? outid : 1];
    idxs.n = outid;
    idxs.d = 1;
    dists.mat = new float[outid > 0 ? outid : 1];
    dists.n = outid;
    dists.d = 1;
    mat res;
    res.mat = const_cast<float *>(residual);
    res.n = nq * w;
    res.d = d;
    preallocate_gpu_mem(PQ, ivf.data(), nlist);
    core_gpu(PQ, res, ivf.data(), nlist, const_cast<int *>(rid_to_ivf), qid_to_starting_outid.data(), idxs, dists, k,
             w);
    deallocate_gpu_mem();
    std::memcpy(out_ids, idxs.mat, sizeof(int) * size_t(outid));
    std::memcpy(out_dis, dists.mat, sizeof(float) * size_t(outid));
    delete[] idxs.mat;
    delete[] dists.mat;
    return outid;
}

// ivfpq_new -> write_cent -> read_cent -> ivfpq_assign.  Outputs the quantizers
// read back from the files and the lists (codes as int32, the reference's).
int64_t run_build(const float *train, int ntrain, const float *base, int nbase, int coarsek, int nsq,
                  const char *fh, const char *fc, const char *fa, float *coarse_out, float *pq_out, int64_t *off,
                  int32_t *ids, int32_t *codes) {
    mat vtrain;
    vtrain.mat = const_cast<float *>(train);
    vtrain.n = ntrain;
    vtrain.d = 128;
    ivfpq_t ivfpq = ivfpq_new(coarsek, nsq, vtrain, 1);
    write_cent(const_cast<char *>(fh), const_cast<char *>(fc), const_cast<char *>(fa), ivfpq);
    ivfpq_t r;
    read_cent(const_cast<char *>(fh), const_cast<char *>(fc), const_cast<char *>(fa), &r);
    if (r.coarsek != coarsek || r.pq.nsq != nsq || r.pq.ks != 256) return -1;
    std::memcpy(coarse_out, r.coa_centroids, sizeof(float) * size_t(coarsek) * 128);
    std::memcpy(pq_out, r.pq.centroids, sizeof(float) * size_t(256) * 128);
    mat vbase;
    vbase.mat = const_cast<float *>(base);
    vbase.n = nbase;
    vbase.d = 128;
    std::vector<ivf_t> ivf(coarsek);
    ivfpq_assign(r, vbase, ivf.data());
    const int64_t n = flatten(ivf.data(), coarsek, nsq, off, ids, codes);
    for (int l = 0; l < coarsek; ++l) {
        std::free(ivf[l].ids);
        std::free(ivf[l].codes.mat);
    }
    std::free(ivfpq.coa_centroids);
    std::free(ivfpq.pq.centroids);
    std::free(r.coa_centroids);
    std::free(r.pq.centroids);
    return n;
}

// read_ivf of the per-list files in dir (genivf layout); returns the vector count.
int64_t run_read_ivf(int coarsek, int nsq, const char *dir, int64_t *off, int32_t *ids, int32_t *codes) {
    ivfpq_t q;
    std::memset(&q, 0, sizeof q);
    q.coarsek = q.coa_centroidsn = coarsek;
    q.pq.nsq = nsq;
    q.pq.ks = 256;
    ivf_t *ivf = read_ivf(q, 0, 0, const_cast<char *>(dir));
    const int64_t n = flatten(ivf, coarsek, nsq, off, ids, codes);
    for (int l = 0; l < coarsek; ++l) {
        std::free(ivf[l].ids);
        std::free(ivf[l].codes.mat);
    }
    std::free(ivf);
    return n;
}

}  // extern "C"
