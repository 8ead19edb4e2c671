// Host side of the batched exact form: per-query w and k, one persistent
// list-major kernel, capture of the whole launch sequence in a hipGraph, and
// an autotuner that picks the work-item shape (queries per item, threads per
// workgroup) on the caller's own batch.
//
// Sequence (all on one stream, no memsets, no host synchronisation, so it can
// be captured): coarse top-wmax (exact) -> probes past each query's w masked
// -> seed lists and thresholds (per-query k) -> list-major pair grouping into
// items of G queries -> exact_scan_kernel -> exact_select_kernel.  After it
// runs, the host reads one overflow count; queries whose candidates exceeded
// the buffer (never seen at the benchmark shapes) are rerun with the plain
// kernels, so the results are always those of mivf_search_var.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/mivf.h"
#include "index_impl.h"

namespace mivf {
namespace {

struct XCfg {
    int g = 8, nt = 512;
};

struct XWs {
    Ws ws;
    DevBuf wq, kq, tau, cnt, poff, ioff, fill, pairq, items, ctrl, cand, ccnt, ovf, runthr, rhist;
    // side stream of the pair grouping (runs beside the seed scan)
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    XWs() = default;
    XWs(const XWs &) = delete;
    XWs &operator=(const XWs &) = delete;
    ~XWs() {
        if (side) (void)hipStreamSynchronize(side);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
    }
    int ensure_side() {
        if (side) return kOk;
        MIVF_CHECK_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        MIVF_CHECK_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        MIVF_CHECK_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
        return kOk;
    }
};

// Candidates per query: the option's capacity, at least 2 kmax, and at least
// batched_cap_pp per probe (the vectors within a query's bound grow with its
// probes).
// Expected candidates per query above which option exact_running_bound = 2
// (automatic) turns the running bounds on: measured at 1,900 (10M vectors,
// 4096 lists, w 32, k 100: 3 % slower with them) and at 4,800 (100M, 16384
// lists, w 32: 5 % faster, and no overflowed query left to rerun).
constexpr double kRunningMinCandidates = 3000.0;

int capacity(const mivf_index *ix, int kmax, int wmax) {
    return std::max(std::max(ix->batched_cap, 2 * kmax), ix->batched_cap_pp * wmax);
}

bool cfg_ok(int m, const XCfg &c) { return exact_lds_bytes(m, c.g) > 0 && (c.nt == 256 || c.nt == 512); }

std::vector<XCfg> candidates(int m) {
    std::vector<XCfg> v;
    const XCfg all[] = {{4, 256}, {4, 512}, {8, 256}, {8, 512}, {16, 512}};
    for (const auto &c : all) {
        // only the compiled shapes
        const int key = (m * 100 + c.g) * 10000 + c.nt;
        if (key == 16040256 || key == 16040512 || key == 16080256 || key == 16080512 || key == 8040256 ||
            key == 8040512 || key == 8080256 || key == 8080512 || key == 8160512)
            v.push_back(c);
    }
    return v;
}

// Enqueue the batched exact search (graph-capturable once the workspace has
// been sized by a first run, with overlap off).  With overlap, the grouping
// of the (list, query) pairs into work items runs on the workspace's side
// stream beside the seed scan: the items do not depend on the seed bounds,
// which the scan reads per query.
int batched_enqueue(mivf_index *ix, XWs &x, const XCfg &cfg, const float *d_q, int nq, int wmax, int kmax,
                    int32_t *d_ids, float *d_dis, hipStream_t st, Timer *tm = nullptr, bool overlap = false) {
    const int nlist = ix->nlist;
    const int capq = capacity(ix, kmax, wmax);
    const int64_t np = int64_t(nq) * wmax;
    MIVF_TRY(x.ws.qn.ensure(sizeof(double) * nq));
    MIVF_TRY(x.ws.dmat.ensure(coarse_ws_bytes(nq, nlist)));
    MIVF_TRY(x.ws.probes.ensure(sizeof(int32_t) * std::max<int64_t>(np, 1)));
    MIVF_TRY(x.ws.coadis.ensure(sizeof(float) * std::max<int64_t>(np, 1)));
    MIVF_TRY(x.tau.ensure(sizeof(float) * nq));
    MIVF_TRY(x.cnt.ensure(sizeof(int) * (nlist + 1)));
    MIVF_TRY(x.poff.ensure(sizeof(int) * (nlist + 1)));
    MIVF_TRY(x.ioff.ensure(sizeof(int) * (nlist + 1)));
    MIVF_TRY(x.fill.ensure(sizeof(int) * (nlist + 1)));
    MIVF_TRY(x.pairq.ensure(sizeof(int32_t) * std::max<int64_t>(np, 1)));
    MIVF_TRY(x.items.ensure(sizeof(FastItem) * std::max<int64_t>(np, 1)));
    MIVF_TRY(x.ctrl.ensure(sizeof(int) * 4));
    MIVF_TRY(x.cand.ensure(sizeof(uint64_t) * size_t(nq) * capq));
    MIVF_TRY(x.ccnt.ensure(sizeof(int) * nq));
    MIVF_TRY(x.ovf.ensure(sizeof(int) * (size_t(nq) + 1)));
    // running bounds pay when many vectors pass the seed bound: about
    // w * (mean list length) * k / seed sample of them per query
    const int seed_cap = ix->exact_seed_cap > 0 ? std::max(ix->exact_seed_cap, kmax) : 0;
    const double expect = seed_cap > 0 ? double(wmax) * double(ix->ntotal) / nlist * kmax / seed_cap : 0.0;
    const bool running = ix->exact_running == 1 || (ix->exact_running == 2 && expect > kRunningMinCandidates);
    if (running) {
        MIVF_TRY(x.runthr.ensure(sizeof(float) * nq));
        MIVF_TRY(x.rhist.ensure(sizeof(int) * size_t(nq) * kExactBins));
    }
    int32_t *probes = x.ws.probes.as<int32_t>();
    int *ctrl = x.ctrl.as<int>();  // [0] total items, [1] next item
    MIVF_TRY(coarse_stage(ix, x.ws, d_q, nq, wmax, probes, x.ws.coadis.as<float>(), st, nullptr));
    if (tm) MIVF_TRY(tm->mark());
    MIVF_CHECK_HIP(launch_mask_probes(probes, nq, wmax, x.wq.as<int32_t>(), st));
    if (overlap) MIVF_TRY(x.ensure_side());
    hipStream_t sp = overlap ? x.side : st;  // stream of the pair grouping
    if (overlap) {
        MIVF_CHECK_HIP(hipEventRecord(x.fork, st));
        MIVF_CHECK_HIP(hipStreamWaitEvent(sp, x.fork, 0));
    }
    {
        // candidate counts, the work counter and the overflow flags start
        // cleared, on the grouping's stream: it follows every earlier search
        // with this workspace (the fork comes after them), and the scan
        // waits for it
        FillSet f{};
        f.p[0] = x.ccnt.as<int>(), f.n[0] = nq, f.v[0] = 0;
        f.p[1] = ctrl + 1, f.n[1] = 1, f.v[1] = 0;
        f.p[2] = x.ovf.as<int>(), f.n[2] = int64_t(nq) + 1, f.v[2] = 0;
        f.count = 3;
        if (running) {  // running bounds (+inf) and their histograms (0)
            f.p[3] = x.runthr.as<int>(), f.n[3] = nq, f.v[3] = 0x7f800000;
            f.p[4] = x.rhist.as<int>(), f.n[4] = int64_t(nq) * kExactBins, f.v[4] = 0;
            f.count = 5;
        }
        MIVF_CHECK_HIP(launch_fill_multi(f, sp));
    }
    MIVF_CHECK_HIP(launch_pairs(probes, nq, wmax, 0, wmax, ix->off.as<int64_t>(), nullptr, nlist, cfg.g,
                                x.cnt.as<int>(), x.poff.as<int>(), x.ioff.as<int>(), ctrl, nullptr,
                                ix->item_order ? ix->lorder.as<int32_t>() : nullptr, x.fill.as<int>(),
                                x.pairq.as<int32_t>(), x.items.as<FastItem>(), sp));
    if (overlap) MIVF_CHECK_HIP(hipEventRecord(x.join, sp));
    SeedArgs sa{};
    sa.queries = d_q;
    sa.coarse = ix->coarse.as<float>();
    sa.pq = ix->pq.as<float>();
    sa.pnorm = ix->pq_n32.as<float>();
    sa.pmax = ix->pmax.as<float>();
    sa.list_off = ix->off.as<int64_t>();
    sa.codes = ix->codes.as<uint8_t>();
    sa.probes = probes;
    sa.min_seed = 1;
    sa.klq = x.kq.as<int32_t>();
    sa.nq = nq;
    sa.w = wmax;
    sa.nlist = nlist;
    sa.ks = ix->ks;
    sa.kl = kmax;
    // the first exact_seed_cap vectors of the seed lists (a prefix stays in
    // LDS in one pass; a longer list costs three passes of table and ADC)
    sa.cap = ix->exact_seed_cap > 0 ? std::max(ix->exact_seed_cap, kmax) : 0;
    sa.tau = x.tau.as<float>();
    MIVF_CHECK_HIP(launch_seed_tau(sa, ix->m, st));
    if (overlap) MIVF_CHECK_HIP(hipStreamWaitEvent(st, x.join, 0));
    ExactArgs a{};
    a.queries = d_q;
    a.coarse = ix->coarse.as<float>();
    a.pq = ix->pq.as<float>();
    a.pq_norm = ix->pq_n64.as<double>();
    a.list_ids = ix->ids.as<int32_t>();
    a.codes = ix->codes.as<uint8_t>();
    a.items = x.items.as<FastItem>();
    a.total_items = ctrl;
    a.next_item = ctrl + 1;
    a.cand = x.cand.as<uint64_t>();
    a.cand_cnt = x.ccnt.as<int>();
    a.capq = capq;
    a.ks = ix->ks;
    a.err = ix->errflag.as<int>();
    a.tau = x.tau.as<float>();
    a.runthr = running ? x.runthr.as<float>() : nullptr;
    a.rhist = running ? x.rhist.as<int>() : nullptr;
    a.kq = x.kq.as<int32_t>();
    a.kuni = kmax;
    const int per_cu = std::max(1, int((160 * 1024) / exact_lds_bytes(ix->m, cfg.g)));
    if (tm) MIVF_TRY(tm->mark());
    MIVF_CHECK_HIP(launch_exact_scan(a, ix->m, cfg.g, cfg.nt, ix->ncu * per_cu, st));
    if (tm) MIVF_TRY(tm->mark());
    MIVF_CHECK_HIP(launch_exact_select(x.cand.as<uint64_t>(), x.ccnt.as<int>(), capq, x.kq.as<int32_t>(), kmax, kmax,
                                       nq, d_ids, d_dis, x.ovf.as<int>(), st));
    if (tm) MIVF_TRY(tm->mark());
    return kOk;
}

// After the batch has finished: rerun overflowed queries with the plain kernels.
int batched_fixup(mivf_index *ix, XWs &x, const float *d_q, int nq, const int32_t *h_w, const int32_t *h_k,
                  int kmax, int32_t *d_ids, float *d_dis, hipStream_t st, int *n_over) {
    int cnt = 0;
    MIVF_CHECK_HIP(hipMemcpyAsync(&cnt, x.ovf.as<int>() + nq, sizeof(int), hipMemcpyDeviceToHost, st));
    MIVF_CHECK_HIP(hipStreamSynchronize(st));
    int err = 0;
    MIVF_CHECK_HIP(hipMemcpy(&err, ix->errflag.p, sizeof(int), hipMemcpyDeviceToHost));
    MIVF_REQUIRE(err == 0, kHip, "batched search: kernel precondition failed (dynamic LDS not at address 0)");
    if (n_over) *n_over = cnt;
    if (cnt == 0) return kOk;
    std::vector<int> flags(nq);
    MIVF_CHECK_HIP(hipMemcpy(flags.data(), x.ovf.p, sizeof(int) * nq, hipMemcpyDeviceToHost));
    std::vector<int> sel;
    for (int q = 0; q < nq; ++q)
        if (flags[q]) sel.push_back(q);
    const int n = int(sel.size());
    const int d = ix->d;
    DevBuf qb, ib, db, wb, kb;
    MIVF_TRY(qb.ensure(sizeof(float) * size_t(n) * d));
    MIVF_TRY(ib.ensure(sizeof(int32_t) * size_t(n) * kmax));
    MIVF_TRY(db.ensure(sizeof(float) * size_t(n) * kmax));
    MIVF_TRY(wb.ensure(sizeof(int32_t) * n));
    MIVF_TRY(kb.ensure(sizeof(int32_t) * n));
    std::vector<int32_t> w(n), k(n);
    int wmax = 1;
    for (int i = 0; i < n; ++i) {
        MIVF_CHECK_HIP(hipMemcpy(qb.as<float>() + size_t(i) * d, d_q + size_t(sel[i]) * d, sizeof(float) * d,
                                 hipMemcpyDeviceToDevice));
        w[i] = h_w[sel[i]];
        k[i] = h_k[sel[i]];
        wmax = std::max(wmax, int(w[i]));
    }
    MIVF_CHECK_HIP(hipMemcpy(wb.p, w.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    MIVF_CHECK_HIP(hipMemcpy(kb.p, k.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    MIVF_TRY(search_entry(ix, x.ws, qb.as<float>(), n, wmax, wb.as<int32_t>(), kb.as<int32_t>(), kmax, kmax,
                          ib.as<int32_t>(), db.as<float>(), st));
    for (int i = 0; i < n; ++i) {
        MIVF_CHECK_HIP(hipMemcpyAsync(d_ids + size_t(sel[i]) * kmax, ib.as<int32_t>() + size_t(i) * kmax,
                                      sizeof(int32_t) * kmax, hipMemcpyDeviceToDevice, st));
        MIVF_CHECK_HIP(hipMemcpyAsync(d_dis + size_t(sel[i]) * kmax, db.as<float>() + size_t(i) * kmax,
                                      sizeof(float) * kmax, hipMemcpyDeviceToDevice, st));
    }
    MIVF_CHECK_HIP(hipStreamSynchronize(st));
    return kOk;
}

int check_wk(const mivf_index *ix, int nq, const int32_t *w, const int32_t *k, int kmax, int *wmax) {
    MIVF_REQUIRE(ix->m == 8 || ix->m == 16, kUnsupported, "batched search: compiled for m in {8, 16}");
    MIVF_REQUIRE(ix->nlist <= kMaxPairLists, kUnsupported, "batched search: nlist above 2^20");
    MIVF_REQUIRE(kmax >= 1 && kmax <= 1024, kInvalid, "batched search: kmax must be in 1..1024");
    *wmax = 1;
    for (int i = 0; i < nq; ++i) {
        MIVF_REQUIRE(w[i] >= 1 && w[i] <= ix->nlist && w[i] <= 1024, kInvalid, "batched search: w[i] out of range");
        MIVF_REQUIRE(k[i] >= 1 && k[i] <= kmax, kInvalid, "batched search: k[i] out of 1..kmax");
        *wmax = std::max(*wmax, int(w[i]));
    }
    return kOk;
}

int upload_wk(XWs &x, int nq, const int32_t *w, const int32_t *k, hipStream_t st) {
    MIVF_TRY(x.wq.ensure(sizeof(int32_t) * nq));
    MIVF_TRY(x.kq.ensure(sizeof(int32_t) * nq));
    MIVF_CHECK_HIP(hipMemcpyAsync(x.wq.p, w, sizeof(int32_t) * nq, hipMemcpyHostToDevice, st));
    MIVF_CHECK_HIP(hipMemcpyAsync(x.kq.p, k, sizeof(int32_t) * nq, hipMemcpyHostToDevice, st));
    return kOk;
}

// Time every compiled work-item shape on this batch; return the fastest.
int autotune(mivf_index *ix, XWs &x, const float *d_q, int nq, int wmax, int kmax, int32_t *d_ids, float *d_dis,
             hipStream_t st, XCfg *best, float *best_ms) {
    hipEvent_t e0, e1;
    MIVF_CHECK_HIP(hipEventCreate(&e0));
    MIVF_CHECK_HIP(hipEventCreate(&e1));
    float bms = 1e30f;
    XCfg b = *best;
    int rc = kOk;
    for (const XCfg &c : candidates(ix->m)) {
        float tot = 0;
        for (int rep = 0; rep < 3 && rc == kOk; ++rep) {  // rep 0 warms up (allocations)
            if (hipEventRecord(e0, st) != hipSuccess) { rc = kHip; break; }
            rc = batched_enqueue(ix, x, c, d_q, nq, wmax, kmax, d_ids, d_dis, st);
            if (rc != kOk) break;
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = kHip; break; }
            float ms = 0;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = kHip; break; }
            if (rep > 0) tot += ms;
        }
        if (rc != kOk) break;
        if (tot / 2 < bms) {
            bms = tot / 2;
            b = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != kOk) return rc;
    *best = b;
    if (best_ms) *best_ms = bms;
    return kOk;
}

XWs &index_xws(mivf_index *ix);

}  // namespace

bool list_major_fits(const mivf_index *ix, int nq) {
    return ix->list_major && nq >= ix->list_major_min && (ix->m == 8 || ix->m == 16) && ix->nlist <= kMaxPairLists &&
           cfg_ok(ix->m, XCfg{ix->batched_g, ix->batched_nt});
}

bool list_major_ok(const mivf_index *ix, const float *d_q, int nq) {
    return list_major_fits(ix, nq) && (reinterpret_cast<uintptr_t>(d_q) & 15u) == 0;
}

// mivf_search of a large batch: the batched list-major kernels with every
// query's w and k equal (filled on the device), overflowed queries rerun by
// the plain kernels.
int search_list_major(mivf_index *ix, const float *d_q, int nq, int w, int k, int32_t *d_ids, float *d_dis,
                      hipStream_t st) {
    MIVF_TRY(upload_lists(ix));
    XWs &x = index_xws(ix);
    const XCfg cfg{ix->batched_g, ix->batched_nt};
    MIVF_TRY(x.wq.ensure(sizeof(int32_t) * nq));
    MIVF_TRY(x.kq.ensure(sizeof(int32_t) * nq));
    {
        FillSet f{};  // every query's w and k, in one launch
        f.p[0] = x.wq.as<int>(), f.n[0] = nq, f.v[0] = w;
        f.p[1] = x.kq.as<int>(), f.n[1] = nq, f.v[1] = k;
        f.count = 2;
        MIVF_CHECK_HIP(launch_fill_multi(f, st));
    }
    Timer tm;
    std::fill(ix->timings, ix->timings + 8, 0.0f);
    MIVF_TRY(tm.start(ix->profile, st, ix->ev));
    MIVF_TRY(batched_enqueue(ix, x, cfg, d_q, nq, w, k, d_ids, d_dis, st, tm.on ? &tm : nullptr, true));
    if (tm.on) {
        MIVF_CHECK_HIP(hipEventSynchronize(tm.ev[tm.n - 1]));
        for (int i = 0; i + 1 < tm.n && i < 8; ++i) {
            float ms = 0;
            MIVF_CHECK_HIP(hipEventElapsedTime(&ms, tm.ev[i], tm.ev[i + 1]));
            ix->timings[i] = ms;
        }
    }
    // the overflow count is read on the host; the per-query w and k of a
    // rerun are only needed when some query overflowed
    float kept[8];
    std::copy(ix->timings, ix->timings + 8, kept);
    const std::vector<int32_t> hw(size_t(nq), w), hk(size_t(nq), k);
    int nov = 0;
    MIVF_TRY(batched_fixup(ix, x, d_q, nq, hw.data(), hk.data(), k, d_ids, d_dis, st, &nov));
    std::copy(kept, kept + 8, ix->timings);
    ix->batched_last_overflow = nov;
    return kOk;
}

std::shared_ptr<void> new_list_major_ws(hipStream_t st) {
    auto x = std::make_shared<XWs>();
    x->ws.order_on(st);
    for (DevBuf *b : {&x->runthr, &x->rhist, &x->wq, &x->kq, &x->tau, &x->cnt, &x->poff, &x->ioff, &x->fill, &x->pairq, &x->items,
                      &x->ctrl, &x->cand, &x->ccnt, &x->ovf})
        b->ord = st;
    return x;
}

// Device bytes per query the list-major kernels hold beyond the plain path's:
// the candidate buffer of capacity(kmax, wmax) (distance, id) pairs.
int64_t list_major_query_bytes(const mivf_index *ix, int kmax, int wmax) {
    return int64_t(capacity(ix, kmax, wmax)) * 8;
}

int list_major_enqueue(mivf_index *ix, void *ws, const float *d_q, int nq, int wmax, const int32_t *d_w,
                       const int32_t *d_k, int kmax, int32_t *d_ids, float *d_dis, int *h_flags, hipStream_t st) {
    XWs &x = *static_cast<XWs *>(ws);
    MIVF_REQUIRE((reinterpret_cast<uintptr_t>(d_q) & 15u) == 0, kInvalid, "list-major batch: queries not 16-byte aligned");
    MIVF_REQUIRE(kmax >= 1 && kmax <= 1024 && wmax >= 1 && wmax <= std::min(ix->nlist, 1024), kInvalid,
                 "list-major batch: w or k out of range");
    MIVF_TRY(x.wq.ensure(sizeof(int32_t) * nq));
    MIVF_TRY(x.kq.ensure(sizeof(int32_t) * nq));
    MIVF_CHECK_HIP(hipMemcpyAsync(x.wq.p, d_w, sizeof(int32_t) * nq, hipMemcpyDeviceToDevice, st));
    MIVF_CHECK_HIP(hipMemcpyAsync(x.kq.p, d_k, sizeof(int32_t) * nq, hipMemcpyDeviceToDevice, st));
    MIVF_TRY(batched_enqueue(ix, x, XCfg{ix->batched_g, ix->batched_nt}, d_q, nq, wmax, kmax, d_ids, d_dis, st,
                             nullptr, true));
    MIVF_CHECK_HIP(hipMemcpyAsync(h_flags, x.ovf.as<int>() + nq, sizeof(int), hipMemcpyDeviceToHost, st));
    MIVF_CHECK_HIP(hipMemcpyAsync(h_flags + 1, ix->errflag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    return kOk;
}

int list_major_finish(mivf_index *ix, void *ws, const float *d_q, int nq, const int32_t *h_w, const int32_t *h_k,
                      int kmax, int32_t *d_ids, float *d_dis, const int *h_flags, hipStream_t st, int *n_over) {
    *n_over = 0;
    MIVF_REQUIRE(h_flags[1] == 0, kHip, "list-major batch: kernel precondition failed (dynamic LDS not at address 0)");
    if (h_flags[0] == 0) return kOk;
    return batched_fixup(ix, *static_cast<XWs *>(ws), d_q, nq, h_w, h_k, kmax, d_ids, d_dis, st, n_over);
}

}  // namespace mivf

using namespace mivf;

struct mivf_plan {
    mivf_index *ix = nullptr;
    XWs x;
    XCfg cfg;
    int nq = 0, wmax = 0, kmax = 0;
    bool tune = false, captured = false;
    uint64_t list_gen = 0;  // the index's list upload the graph was captured against
    DevBuf q, ids, dis;
    int32_t *h_wk = nullptr;  // pinned staging of w and k
    hipStream_t st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float tuned_ms = 0;
    ~mivf_plan() {
        if (st) (void)hipStreamSynchronize(st);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (st) (void)hipStreamDestroy(st);
        if (h_wk) (void)hipHostFree(h_wk);
    }
};

namespace mivf {
namespace {
XWs &index_xws(mivf_index *ix) {
    if (!ix->xws) ix->xws = std::make_shared<XWs>();
    return *std::static_pointer_cast<XWs>(ix->xws);
}
}  // namespace
}  // namespace mivf

int mivf_search_batched(mivf_index *ix, const float *d_q, int nq, const int32_t *w, const int32_t *k, int kmax,
                        int32_t *d_ids, float *d_dis, void *stream) try {
    MIVF_REQUIRE(ix, kInvalid, "null index");
    if (nq <= 0) return kOk;
    MIVF_REQUIRE(d_q && w && k && d_ids && d_dis, kInvalid, "search_batched: null pointer");
    MIVF_REQUIRE((reinterpret_cast<uintptr_t>(d_q) & 15u) == 0, kInvalid, "search_batched: d_q must be 16-byte aligned");
    int wmax = 1;
    MIVF_TRY(check_wk(ix, nq, w, k, kmax, &wmax));
    MIVF_TRY(upload_lists(ix));
    hipStream_t st = static_cast<hipStream_t>(stream);
    XWs &x = index_xws(ix);
    XCfg cfg;
    cfg.g = ix->batched_g;
    cfg.nt = ix->batched_nt;
    MIVF_REQUIRE(cfg_ok(ix->m, cfg), kInvalid, "search_batched: work-item shape not compiled");
    MIVF_TRY(upload_wk(x, nq, w, k, st));
    Timer tm;
    std::fill(ix->timings, ix->timings + 8, 0.0f);
    MIVF_TRY(tm.start(ix->profile, st, ix->ev));
    MIVF_TRY(batched_enqueue(ix, x, cfg, d_q, nq, wmax, kmax, d_ids, d_dis, st, tm.on ? &tm : nullptr, true));
    if (tm.on) {
        MIVF_CHECK_HIP(hipEventSynchronize(tm.ev[tm.n - 1]));
        for (int i = 0; i + 1 < tm.n && i < 8; ++i) {
            float ms = 0;
            MIVF_CHECK_HIP(hipEventElapsedTime(&ms, tm.ev[i], tm.ev[i + 1]));
            ix->timings[i] = ms;
        }
    }
    int nov = 0;
    float kept[8];
    std::copy(ix->timings, ix->timings + 8, kept);  // the rerun of overflowed queries records its own
    MIVF_TRY(batched_fixup(ix, x, d_q, nq, w, k, kmax, d_ids, d_dis, st, &nov));
    std::copy(kept, kept + 8, ix->timings);
    ix->batched_last_overflow = nov;
    return kOk;
} MIVF_CATCH_ALL

int mivf_batched_stats(mivf_index *ix, int64_t *overflowed_queries) try {
    MIVF_REQUIRE(ix, kInvalid, "null index");
    if (overflowed_queries) *overflowed_queries = ix->batched_last_overflow;
    return kOk;
} MIVF_CATCH_ALL

int mivf_plan_create(mivf_index *ix, int nq, int wmax, int kmax, int autotune_flag, mivf_plan **out) try {
    MIVF_REQUIRE(ix && out && nq >= 1, kInvalid, "plan_create: bad arguments");
    MIVF_REQUIRE(wmax >= 1 && wmax <= ix->nlist && wmax <= 1024 && kmax >= 1 && kmax <= 1024, kInvalid,
                 "plan_create: wmax must be in 1..min(nlist,1024), kmax in 1..1024");
    MIVF_REQUIRE(ix->m == 8 || ix->m == 16, kUnsupported, "plan_create: compiled for m in {8, 16}");
    MIVF_TRY(upload_lists(ix));
    auto *p = new mivf_plan();
    p->ix = ix;
    p->nq = nq;
    p->wmax = wmax;
    p->kmax = kmax;
    p->tune = autotune_flag != 0;
    p->cfg.g = ix->batched_g;
    p->cfg.nt = ix->batched_nt;
    int rc = kOk;
    if (hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&p->h_wk), sizeof(int32_t) * 2 * size_t(nq)) != hipSuccess) {
        set_error("plan_create: stream or pinned buffer");
        rc = kHip;
    }
    if (rc == kOk) rc = p->q.ensure(sizeof(float) * size_t(nq) * ix->d);
    if (rc == kOk) rc = p->ids.ensure(sizeof(int32_t) * size_t(nq) * kmax);
    if (rc == kOk) rc = p->dis.ensure(sizeof(float) * size_t(nq) * kmax);
    if (rc == kOk) rc = p->x.wq.ensure(sizeof(int32_t) * nq);
    if (rc == kOk) rc = p->x.kq.ensure(sizeof(int32_t) * nq);
    if (rc != kOk) {
        delete p;
        return rc;
    }
    *out = p;
    return kOk;
} MIVF_CATCH_ALL

int mivf_plan_free(mivf_plan *p) try {
    delete p;
    return kOk;
} MIVF_CATCH_ALL

int mivf_plan_search(mivf_plan *p, const float *d_q, const int32_t *w, const int32_t *k, int32_t *d_ids, float *d_dis,
                     void *stream) try {
    MIVF_REQUIRE(p && d_q && w && k && d_ids && d_dis, kInvalid, "plan_search: null argument");
    mivf_index *ix = p->ix;
    const int nq = p->nq;
    for (int i = 0; i < nq; ++i) {
        MIVF_REQUIRE(w[i] >= 1 && w[i] <= p->wmax && k[i] >= 1 && k[i] <= p->kmax, kInvalid,
                     "plan_search: w[i] or k[i] outside the plan's wmax / kmax");
        p->h_wk[i] = w[i];
        p->h_wk[nq + i] = k[i];
    }
    hipStream_t user = static_cast<hipStream_t>(stream);
    hipStream_t st = p->st;
    if (user) {  // order after the caller's work that produced d_q
        hipEvent_t e;
        MIVF_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        MIVF_CHECK_HIP(hipEventRecord(e, user));
        MIVF_CHECK_HIP(hipStreamWaitEvent(st, e, 0));
        MIVF_CHECK_HIP(hipEventDestroy(e));
    }
    // The graph holds the device addresses of the lists: after an add or
    // set_lists, upload the new lists and capture again.
    MIVF_TRY(upload_lists(ix));
    if (p->captured && p->list_gen != ix->list_gen) {
        MIVF_CHECK_HIP(hipStreamSynchronize(st));
        MIVF_CHECK_HIP(hipGraphExecDestroy(p->exec));
        MIVF_CHECK_HIP(hipGraphDestroy(p->graph));
        p->exec = nullptr;
        p->graph = nullptr;
        p->captured = false;
        p->tune = false;  // keep the tuned work-item shape
    }
    MIVF_CHECK_HIP(hipMemcpyAsync(p->q.p, d_q, sizeof(float) * size_t(nq) * ix->d, hipMemcpyDeviceToDevice, st));
    MIVF_CHECK_HIP(hipMemcpyAsync(p->x.wq.p, p->h_wk, sizeof(int32_t) * nq, hipMemcpyHostToDevice, st));
    MIVF_CHECK_HIP(hipMemcpyAsync(p->x.kq.p, p->h_wk + nq, sizeof(int32_t) * nq, hipMemcpyHostToDevice, st));
    if (!p->captured) {
        if (p->tune)
            MIVF_TRY(autotune(ix, p->x, p->q.as<float>(), nq, p->wmax, p->kmax, p->ids.as<int32_t>(),
                              p->dis.as<float>(), st, &p->cfg, &p->tuned_ms));
        // eager run: sizes every buffer and checks the kernels' preconditions
        MIVF_TRY(batched_enqueue(ix, p->x, p->cfg, p->q.as<float>(), nq, p->wmax, p->kmax, p->ids.as<int32_t>(),
                                 p->dis.as<float>(), st));
        MIVF_CHECK_HIP(hipStreamSynchronize(st));
        MIVF_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = batched_enqueue(ix, p->x, p->cfg, p->q.as<float>(), nq, p->wmax, p->kmax, p->ids.as<int32_t>(),
                                       p->dis.as<float>(), st);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(st, &g);
        MIVF_TRY(rc);
        MIVF_CHECK_HIP(ce);
        p->graph = g;
        MIVF_CHECK_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
        p->captured = true;
        p->list_gen = ix->list_gen;
    }
    MIVF_CHECK_HIP(hipGraphLaunch(p->exec, st));
    int nov = 0;
    MIVF_TRY(batched_fixup(ix, p->x, p->q.as<float>(), nq, w, k, p->kmax, p->ids.as<int32_t>(), p->dis.as<float>(), st,
                           &nov));
    ix->batched_last_overflow = nov;
    MIVF_CHECK_HIP(hipMemcpyAsync(d_ids, p->ids.p, sizeof(int32_t) * size_t(nq) * p->kmax, hipMemcpyDeviceToDevice, st));
    MIVF_CHECK_HIP(hipMemcpyAsync(d_dis, p->dis.p, sizeof(float) * size_t(nq) * p->kmax, hipMemcpyDeviceToDevice, st));
    MIVF_CHECK_HIP(hipStreamSynchronize(st));
    return kOk;
} MIVF_CATCH_ALL

int mivf_plan_info(const mivf_plan *p, int *g, int *threads, float *tuned_ms) try {
    MIVF_REQUIRE(p, kInvalid, "plan_info: null plan");
    if (g) *g = p->cfg.g;
    if (threads) *threads = p->cfg.nt;
    if (tuned_ms) *tuned_ms = p->tuned_ms;
    return kOk;
} MIVF_CATCH_ALL

int mivf_exact_tables(mivf_index *ix, const float *d_q, int n, int list, float *d_out) try {
    MIVF_REQUIRE(ix && d_q && d_out && n >= 1 && n <= 8 && list >= 0 && list < ix->nlist, kInvalid,
                 "exact_tables: bad arguments");
    ExactArgs a{};
    a.queries = d_q;
    a.coarse = ix->coarse.as<float>();
    a.pq = ix->pq.as<float>();
    a.pq_norm = ix->pq_n64.as<double>();
    a.ks = ix->ks;
    MIVF_CHECK_HIP(launch_exact_tables(a, ix->m, list, n, d_out, nullptr));
    MIVF_CHECK_HIP(hipDeviceSynchronize());
    return kOk;
} MIVF_CATCH_ALL

int mivf_exact_adc(mivf_index *ix, const float *d_q, int n, int list, float *d_out) try {
    MIVF_REQUIRE(ix && d_q && d_out && n >= 1 && n <= 8 && list >= 0 && list < ix->nlist, kInvalid,
                 "exact_adc: bad arguments");
    MIVF_REQUIRE(ix->m == 16, kUnsupported, "exact_adc: compiled for m = 16");
    MIVF_TRY(upload_lists(ix));
    ExactArgs a{};
    a.queries = d_q;
    a.coarse = ix->coarse.as<float>();
    a.pq = ix->pq.as<float>();
    a.pq_norm = ix->pq_n64.as<double>();
    a.codes = ix->codes.as<uint8_t>();
    a.ks = ix->ks;
    const int64_t beg = ix->h_off[list];
    const int len = int(ix->h_off[list + 1] - beg);
    MIVF_CHECK_HIP(launch_exact_adc(a, ix->m, list, n, beg, len, d_out, nullptr));
    MIVF_CHECK_HIP(hipDeviceSynchronize());
    return kOk;
} MIVF_CATCH_ALL

