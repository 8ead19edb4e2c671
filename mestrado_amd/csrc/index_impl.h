// Host-side internals shared by the index layer (index.hip) and the serving
// loop (serve.hip): device buffers, stage timers, search workspaces and the
// index handle.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "mivf_internal.h"

namespace mivf {

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    // When set before the first allocation: the only stream that uses the
    // buffer; it then grows stream-ordered (hipFreeAsync / hipMallocAsync),
    // without the device-wide wait of hipFree.
    hipStream_t ord = nullptr;
    bool async = false;
    // A view into memory the buffer does not own (a slice of a fixed scratch
    // array): it never allocates or frees, and a request beyond its size fails.
    bool fixed = false;
    void view(void *at, size_t n) {
        p = at;
        bytes = n;
        fixed = true;
    }
    int ensure(size_t want) {
        if (want <= bytes) return kOk;
        if (fixed) {
            set_error("scratch view of " + std::to_string(bytes) + " bytes, " + std::to_string(want) + " needed");
            return kNoMem;
        }
        if (p) {
            if (async)
                (void)hipFreeAsync(p, ord);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        async = ord != nullptr;
        hipError_t e = async ? hipMallocAsync(&p, want, ord) : hipMalloc(&p, want);
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc(") + std::to_string(want) + "): " + hipGetErrorString(e));
            return kNoMem;
        }
        bytes = want;
        return kOk;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
    void release() {
        if (p && !fixed) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~DevBuf() { release(); }
};

// Per-stage HIP event timing on the launch stream (events owned by the index).
struct Timer {
    bool on = false;
    hipStream_t s = nullptr;
    hipEvent_t *ev = nullptr;
    int n = 0;
    int start(bool enable, hipStream_t st, hipEvent_t *evs) {
        on = enable && evs;
        s = st;
        n = 0;
        ev = evs;
        if (!on) return kOk;
        for (int i = 0; i < 8; ++i)
            if (!ev[i]) MIVF_CHECK_HIP(hipEventCreate(&ev[i]));
        return mark();
    }
    int mark() {
        if (!on) return kOk;
        MIVF_CHECK_HIP(hipEventRecord(ev[n++], s));
        return kOk;
    }
};

// Scratch of one search in flight (grow-only).  The index owns one; each
// serving slot owns another, so searches on different streams can overlap.
struct Ws {
    DevBuf qn, dmat, probes, coadis, partial, wq, kq, redo, tmp_ids, tmp_dis, qorder, rres, rnd;
    // every buffer grows stream-ordered on s (the only stream these calls use)
    void order_on(hipStream_t s) {
        for (DevBuf *b : all()) b->ord = s;
    }
    void release() {
        for (DevBuf *b : all()) b->release();
    }
    std::array<DevBuf *, 13> all() {
        return {&qn, &dmat, &probes, &coadis, &partial, &wq, &kq, &redo, &tmp_ids, &tmp_dis, &qorder, &rres, &rnd};
    }
};

}  // namespace mivf

using namespace mivf;

struct mivf_index {
    int d = 0, nlist = 0, m = 0, ks = 0, ds = 0;
    // quantizers (device)
    DevBuf coarse, coarse_n32, pq, pq_n64, pq_n32;
    float maxcn = 0.0f;      // max_c |c|^2 (margin of the coarse recheck)
    DevBuf coarse_hi, coarse_lo;  // BF16 hi/lo split of the centroids (matrix-core coarse distances)
    // lists: host mirror (list-major) + device arrays
    std::vector<std::vector<int32_t>> h_ids;
    std::vector<std::vector<uint8_t>> h_codes;
    bool dirty = true;
    uint64_t list_gen = 0;   // bumped by every upload of the lists (captured graphs compare it)
    int64_t ntotal = 0;
    std::vector<int64_t> h_off;
    DevBuf off, ids, codes;
    // scratch
    Ws ws;                   // scratch of the calls made on this handle
    DevBuf keys, res, resn, assign;
    DevBuf pmax;             // [m] max_c |p_sc| (reduced-precision table bound)
    DevBuf fpb;              // fp16 B operands of the reduced-precision tables
    DevBuf rows;             // raw vectors in list order (mivf_index_attach_base)
    bool rows_ok = false;
    bool rows_u8 = false;    // rows holds bytes (option base_bytes: integer-valued data 0..255)
    bool base_bytes = false; // option base_bytes: attach_base keeps a byte copy
    int fast_front = 0;      // option fast_front_ranks: probe ranks whose work items the 8-bit scan takes first
    int item_order = 0;      // option item_order: 1 = list-major work items numbered longest list first
    int exact_running = 2;   // option exact_running_bound: running per-query bounds in the exact list-major scan (2: automatic)
    DevBuf lorder;           // [nlist] lists by decreasing length (set by upload_lists)
    // batched reduced-precision form scratch
    DevBuf f_tau, f_cnt, f_poff, f_ioff, f_fill, f_pairq, f_items, f_ctrl, f_cand, f_ccnt,
        f_short, f_rthr, f_rhist;
    int64_t f_last_cands = 0, f_last_overflow = 0;
    DevBuf errflag;          // kernel-reported precondition failures (int)
    bool scan_checked = false;
    // options
    int ppg_opt = 0;
    int query_order = 1;      // scan queries grouped by nearest list, per XCD (scheduling only)
    bool fast_running = true;  // running per-query bound in the reduced-precision scan
    int fast_seed = 1;        // lists scanned exactly to seed the reduced-precision filter
    int fast_seed_cap = 2048; // vectors entering the seed bound (0 = all of the seed lists)
    int fast_slack = 2;       // option fast_seed_slack: quanta above the seed bound that pass the 8-bit filter
    int exact_seed_cap = 4096; // the same for the list-major exact kernels
    int fast_cap = 65536;     // candidate capacity per query of the reduced-precision form
    int batched_cap = 16384;  // candidate capacity per query of the batched exact form (>= 2 kmax)
    int batched_cap_pp = 256; // ... and at least this many per probe
    int batched_g = 4, batched_nt = 512;  // its work-item shape (queries per item, threads)
    // mivf_search on batches of at least list_major_min queries runs the
    // batched list-major kernels (identical results)
    bool list_major = true;
    // coarse top-w through the BF16 matrix-core distances + exact recheck
    // (0: the exact FP32 distance tiles + row selection for every shape)
    bool coarse_mfma = true;
    // table type of mivf_search_fast: 0 = 8-bit entries built on the FP16
    // matrix cores, one scale per (query, list), int8 matrix-core sums;
    // 1 = FP16 entries, 2 = 8-bit entries with a per-subquantizer offset and
    // scale (both built from FP32 tables, FP32 sums)
    int fast_tables = 0;
    int64_t list_major_min = 2048;
    // serving: while a micro-batch is in flight, a free slot waits until at
    // least serve_min_fill queries wait or the oldest has waited
    // serve_max_wait_us (min_fill 0: launch whatever waits as soon as a slot
    // is free)
    int64_t serve_min_fill = 1024, serve_max_wait_us = 300;
    int batched_last_overflow = 0;
    int ncu = 256;            // compute units of the device
    std::shared_ptr<void> xws;  // workspace of the batched exact form (batched.hip)
    bool profile = false;
    float timings[8] = {};
    hipEvent_t ev[8] = {};
    // second stream of the reduced-precision form: the (list, query) grouping
    // runs there while the seed kernel runs on the caller's stream
    hipStream_t aux = nullptr;
    hipEvent_t aux_ev[2] = {};
    ~mivf_index() {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : aux_ev)
            if (e) (void)hipEventDestroy(e);
        if (aux) (void)hipStreamDestroy(aux);
    }
};


namespace mivf {

// Search over nq device-resident queries with the handle's kernels and the
// given workspace, on stream st.  d_wq/d_kq: per-query w and k (device) or
// null for the uniform wmax / kuni.  Outputs [nq][kstride].
int search_entry(mivf_index *ix, Ws &ws, const float *d_q, int nq, int wmax, const int32_t *d_wq,
                 const int32_t *d_kq, int kuni, int kstride, int32_t *d_ids, float *d_dis, hipStream_t st);
int upload_lists(mivf_index *ix);
// Records [row0, row0 + n) of an fvecs / ivecs (elem 4) or bvecs (elem 1)
// file of dimension d, converted to T (io.hip).
template <typename T>
int read_vecs_rows(const char *path, int elem, int64_t row0, int64_t n, int d, T *out);
// Whether mivf_search sends this batch to the list-major kernels, and that
// path (batched.hip): uniform w and k, results identical to search_entry.
bool list_major_ok(const mivf_index *ix, const float *d_q, int nq);
int search_list_major(mivf_index *ix, const float *d_q, int nq, int w, int k, int32_t *d_ids, float *d_dis,
                      hipStream_t st);
// One serving micro-batch through the list-major kernels on the caller's
// stream and workspace (batched.hip).  list_major_enqueue leaves the
// overflow count and the kernel error flag in pinned h_flags[0..1] when the
// stream reaches them; list_major_finish, called once the stream has passed
// that point, reruns overflowed queries with the plain kernels (n_over > 0:
// rows of d_ids/d_dis were rewritten after the enqueued work).
std::shared_ptr<void> new_list_major_ws(hipStream_t st);
bool list_major_fits(const mivf_index *ix, int nq);
int64_t list_major_query_bytes(const mivf_index *ix, int kmax, int wmax);
int list_major_enqueue(mivf_index *ix, void *ws, const float *d_q, int nq, int wmax, const int32_t *d_w,
                       const int32_t *d_k, int kmax, int32_t *d_ids, float *d_dis, int *h_flags, hipStream_t st);
int list_major_finish(mivf_index *ix, void *ws, const float *d_q, int nq, const int32_t *h_w, const int32_t *h_k,
                      int kmax, int32_t *d_ids, float *d_dis, const int *h_flags, hipStream_t st, int *n_over);
// ws.dmat of a coarse stage of nq queries: the distance matrix [nq][nlist]
// (offset 0), the 16-centroid group minima [nq][ceil(nlist / 16)] and the
// BF16 hi / lo terms of the queries [nq][128] each (byte offsets, 256-aligned).
struct CoarseWs {
    size_t gmin, qhi, qlo, bytes;
};
inline CoarseWs coarse_ws(int64_t nq, int nlist) {
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    CoarseWs l;
    l.gmin = up(sizeof(float) * size_t(nq) * size_t(nlist));
    l.qhi = up(l.gmin + sizeof(float) * size_t(nq) * size_t((nlist + 15) / 16));
    l.qlo = up(l.qhi + 2 * size_t(nq) * 128);
    l.bytes = up(l.qlo + 2 * size_t(nq) * 128);
    return l;
}
inline size_t coarse_ws_bytes(int64_t nq, int nlist) { return coarse_ws(nq, nlist).bytes; }
// Coarse top-w of nq queries into probes/dis [nq][w] (ws.dmat sized coarse_ws_bytes).
int coarse_stage(mivf_index *ix, Ws &ws, const float *d_q, int nq, int w, int32_t *probes, float *dis, hipStream_t st,
                 Timer *tm);

}  // namespace mivf
