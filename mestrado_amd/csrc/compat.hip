// Reference-signature entry points of the GPU fork (parallel2/) on top of the
// C ABI: the device boundary core_gpu / preallocate_gpu_mem /
// deallocate_gpu_mem (parallel2/ivf_pq/mycuda.h:7-9) and the index build and
// load calls of parallel2/ivf_pq/ivf_training.h and ivf_search.h.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/mestrado_compat.h"
#include "../../include/mivf.h"

namespace {

// Device state is one process-wide handle, as in the reference
// (static gpu_PQ / gpu_ivf, parallel2/ivf_pq/mycuda.cu:21-22).
mivf_index *g_ix = nullptr;
std::vector<int> g_sizes;

[[noreturn]] void die(const char *call) {
    std::fprintf(stderr, "Failed call: %s\nError: %s\n", call, mivf_last_error());
    std::exit(EXIT_FAILURE);
}
#define COMPAT(call)                  \
    do {                              \
        if ((call) != MIVF_OK) die(#call); \
    } while (0)
#define COMPAT_REQUIRE(cond, msg)                                  \
    do {                                                           \
        if (!(cond)) {                                             \
            std::fprintf(stderr, "Invalid argument: %s\n", msg);   \
            std::exit(EXIT_FAILURE);                               \
        }                                                          \
    } while (0)

struct DevMem {
    void *p = nullptr;
    explicit DevMem(size_t bytes) { COMPAT(mivf_malloc(&p, bytes)); }
    ~DevMem() { mivf_free(p); }
};

void lists_to_flat(const ivf_t *ivf, int nlist, int m, std::vector<int64_t> &off, std::vector<int32_t> &ids,
                   std::vector<int32_t> &codes) {
    off.assign(nlist + 1, 0);
    for (int l = 0; l < nlist; ++l) {
        COMPAT_REQUIRE(ivf[l].idstam >= 0 && ivf[l].codes.n == ivf[l].idstam, "ivf_t: codes.n != idstam");
        COMPAT_REQUIRE(ivf[l].idstam == 0 || ivf[l].codes.d == m, "ivf_t: codes.d != nsq");
        off[l + 1] = off[l] + ivf[l].idstam;
    }
    ids.resize(off[nlist]);
    codes.resize(size_t(off[nlist]) * m);
    for (int l = 0; l < nlist; ++l) {
        if (!ivf[l].idstam) continue;
        std::memcpy(ids.data() + off[l], ivf[l].ids, sizeof(int32_t) * ivf[l].idstam);
        std::memcpy(codes.data() + off[l] * m, ivf[l].codes.mat, sizeof(int32_t) * size_t(ivf[l].idstam) * m);
    }
}

void flat_to_lists(const std::vector<int64_t> &off, const int32_t *ids, const uint8_t *codes, int nlist, int m,
                   ivf_t *ivf) {
    for (int l = 0; l < nlist; ++l) {
        const int64_t n = off[l + 1] - off[l];
        ivf[l].idstam = int(n);
        ivf[l].ids = static_cast<int *>(std::malloc(sizeof(int) * (n ? n : 1)));
        ivf[l].codes.n = int(n);
        ivf[l].codes.d = m;
        ivf[l].codes.mat = static_cast<int *>(std::malloc(sizeof(int) * (n ? n * m : 1)));
        for (int64_t i = 0; i < n; ++i) ivf[l].ids[i] = ids[off[l] + i];
        for (int64_t i = 0; i < n * m; ++i) ivf[l].codes.mat[i] = codes[off[l] * m + i];
    }
}

}  // namespace

void preallocate_gpu_mem(pqtipo PQ, ivf_t *ivf, int ivf_size) {
    if (g_ix) deallocate_gpu_mem();
    COMPAT_REQUIRE(ivf_size > 0 && PQ.nsq > 0 && PQ.ds > 0 && PQ.centroids, "preallocate_gpu_mem");
    const int d = PQ.nsq * PQ.ds;
    // the device boundary receives residuals, so the coarse centroids are not needed here
    std::vector<float> zeros(size_t(ivf_size) * d, 0.0f);
    COMPAT(mivf_index_create(d, ivf_size, PQ.nsq, PQ.ks, zeros.data(), PQ.centroids, &g_ix));
    std::vector<int64_t> off;
    std::vector<int32_t> ids, codes;
    lists_to_flat(ivf, ivf_size, PQ.nsq, off, ids, codes);
    COMPAT(mivf_index_set_lists_i32(g_ix, off.data(), ids.data(), codes.data()));
    g_sizes.resize(ivf_size);
    for (int l = 0; l < ivf_size; ++l) g_sizes[l] = ivf[l].idstam;
}

void deallocate_gpu_mem() {
    if (g_ix) mivf_index_free(g_ix);
    g_ix = nullptr;
    g_sizes.clear();
}

void core_gpu(pqtipo PQ, mat residual, ivf_t *ivf, int ivf_size, int *rid_to_ivf, int *qid_to_starting_outid,
              matI idxs, mat dists, int k, int w) {
    (void)PQ;
    (void)ivf;
    COMPAT_REQUIRE(g_ix, "core_gpu called before preallocate_gpu_mem");
    COMPAT_REQUIRE(w > 0 && k > 0 && residual.n % w == 0, "core_gpu: residual.n must be nq * w");
    COMPAT_REQUIRE(ivf_size == int(g_sizes.size()), "core_gpu: ivf_size differs from preallocate_gpu_mem");
    const int nq = residual.n / w;
    if (nq == 0) return;
    for (int i = 0; i < residual.n; ++i)
        COMPAT_REQUIRE(rid_to_ivf[i] >= 0 && rid_to_ivf[i] < ivf_size, "core_gpu: rid_to_ivf out of range");
    const size_t rb = sizeof(float) * size_t(residual.n) * residual.d;
    DevMem dres(rb), dlist(sizeof(int) * residual.n), dids(sizeof(int) * size_t(nq) * k),
        ddis(sizeof(float) * size_t(nq) * k);
    COMPAT(mivf_memcpy_htod(dres.p, residual.mat, rb));
    COMPAT(mivf_memcpy_htod(dlist.p, rid_to_ivf, sizeof(int) * residual.n));
    COMPAT(mivf_search_residual(g_ix, static_cast<float *>(dres.p), static_cast<int32_t *>(dlist.p), nq, w, k,
                                static_cast<int32_t *>(dids.p), static_cast<float *>(ddis.p), nullptr));
    std::vector<int32_t> hid(size_t(nq) * k);
    std::vector<float> hdis(size_t(nq) * k);
    COMPAT(mivf_memcpy_dtoh(hid.data(), dids.p, sizeof(int32_t) * hid.size()));
    COMPAT(mivf_memcpy_dtoh(hdis.data(), ddis.p, sizeof(float) * hdis.size()));
    for (int q = 0; q < nq; ++q) {
        int64_t n = 0;
        for (int p = 0; p < w; ++p) n += g_sizes[rid_to_ivf[q * w + p]];
        const int cnt = int(n < k ? n : k);
        const int out = qid_to_starting_outid[q];
        COMPAT_REQUIRE(out >= 0 && out + cnt <= idxs.n && out + cnt <= dists.n, "core_gpu: output too small");
        for (int i = 0; i < cnt; ++i) {
            idxs.mat[out + i] = hid[size_t(q) * k + i];
            dists.mat[out + i] = hdis[size_t(q) * k + i];
        }
    }
}

ivfpq_t ivfpq_new(int coarsek, int nsq, mat vtrain, int threads) {
    (void)threads;
    COMPAT_REQUIRE(coarsek > 0 && nsq > 0 && vtrain.d % nsq == 0, "ivfpq_new");
    const int d = vtrain.d, ks = 256, ds = d / nsq;
    ivfpq_t r;
    r.coarsek = r.coa_centroidsn = coarsek;
    r.coa_centroidsd = d;
    r.coa_centroids = static_cast<float *>(std::malloc(sizeof(float) * size_t(coarsek) * d));
    r.pq.nsq = nsq;
    r.pq.ks = ks;
    r.pq.ds = ds;
    r.pq.centroidsn = ks * ds;
    r.pq.centroidsd = nsq;
    r.pq.centroids = static_cast<float *>(std::malloc(sizeof(float) * size_t(ks) * d));
    // iteration counts and seed of ivfpq_new / pq_new (ivf_training.cpp:77, pq_new.cpp:111)
    COMPAT(mivf_train(vtrain.mat, vtrain.n, d, coarsek, nsq, ks, 50, 100, 2, r.coa_centroids, r.pq.centroids));
    return r;
}

void ivfpq_assign(ivfpq_t ivfpq, mat vbase, ivf_t *ivf) {
    mivf_index *ix = nullptr;
    COMPAT(mivf_index_create(ivfpq.coa_centroidsd, ivfpq.coarsek, ivfpq.pq.nsq, ivfpq.pq.ks, ivfpq.coa_centroids,
                             ivfpq.pq.centroids, &ix));
    if (mivf_index_add(ix, vbase.mat, vbase.n, 0) != MIVF_OK) die("mivf_index_add");
    std::vector<int64_t> off(ivfpq.coarsek + 1);
    std::vector<int32_t> ids(vbase.n > 0 ? vbase.n : 1);
    std::vector<uint8_t> codes(size_t(vbase.n > 0 ? vbase.n : 1) * ivfpq.pq.nsq);
    if (mivf_index_get_lists(ix, off.data(), ids.data(), codes.data()) != MIVF_OK) die("mivf_index_get_lists");
    mivf_index_free(ix);
    flat_to_lists(off, ids.data(), codes.data(), ivfpq.coarsek, ivfpq.pq.nsq, ivf);
}

void write_cent(char *file, char *file2, char *file3, ivfpq_t ivfpq) {
    COMPAT(mivf_write_cent(file, file2, file3, ivfpq.pq.nsq, ivfpq.pq.ks, ivfpq.pq.ds, ivfpq.coarsek,
                           ivfpq.coa_centroidsd, ivfpq.pq.centroids, ivfpq.coa_centroids));
}

void read_cent(char *file, char *file2, char *file3, ivfpq_t *ivfpq) {
    int nsq, ks, ds, coarsek, d;
    COMPAT(mivf_read_cent_header(file, &nsq, &ks, &ds, &coarsek, &d));
    ivfpq->pq.nsq = nsq;
    ivfpq->pq.ks = ks;
    ivfpq->pq.ds = ds;
    ivfpq->pq.centroidsn = ks * ds;
    ivfpq->pq.centroidsd = nsq;
    ivfpq->coarsek = ivfpq->coa_centroidsn = coarsek;
    ivfpq->coa_centroidsd = d;
    ivfpq->pq.centroids = static_cast<float *>(std::malloc(sizeof(float) * size_t(nsq) * ks * ds));
    ivfpq->coa_centroids = static_cast<float *>(std::malloc(sizeof(float) * size_t(coarsek) * d));
    COMPAT(mivf_read_cent(file, file2, file3, ivfpq->pq.centroids, ivfpq->coa_centroids));
}

ivf_t *read_ivf(ivfpq_t ivfpq, int tam, int my_rank, char *ivf_path) {
    (void)tam;
    (void)my_rank;
    const int nl = ivfpq.coarsek, m = ivfpq.pq.nsq;
    std::vector<int64_t> off(nl + 1);
    COMPAT(mivf_read_ivf_dir(ivf_path, nl, m, off.data(), nullptr, nullptr, 0));
    std::vector<int32_t> ids(off[nl] > 0 ? off[nl] : 1);
    std::vector<uint8_t> codes(size_t(off[nl] > 0 ? off[nl] : 1) * m);
    COMPAT(mivf_read_ivf_dir(ivf_path, nl, m, off.data(), ids.data(), codes.data(), off[nl]));
    ivf_t *ivf = static_cast<ivf_t *>(std::malloc(sizeof(ivf_t) * nl));
    flat_to_lists(off, ids.data(), codes.data(), nl, m, ivf);
    return ivf;
}
