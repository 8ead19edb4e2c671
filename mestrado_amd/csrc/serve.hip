// Serving loop: continuous batching of search requests over one index.
//
// A request is a batch of queries already in HBM (d_q [nq][d]) with its own
// w and k, and device outputs [nq][k].  The server keeps `slots` micro-batches
// in flight, each on its own HIP stream with its own search workspace.
// Whenever a slot is free, the requests waiting in arrival order are packed
// into it (their queries copied next to each other, per-query w and k
// uploaded) as long as the micro-batch stays within max_batch queries and
// the admitted per-request device state stays within the memory budget; the
// slot runs the exact search with per-query w and k and copies each
// request's rows to its outputs.  A micro-batch of at least the index's
// list-major threshold (the option list_major_min_queries, as mivf_search
// uses it) goes through the list-major kernels of mivf_search_batched, a
// smaller one through the per-query kernels of mivf_search_var; both give
// the same rows bit for bit, so a request's results depend only on its own
// queries, w and k and are identical alone and in any mix.
//
// The reference serves queries through MPI roles (a query sender, search
// ranks and an aggregator, parallel/ivf_test.cpp); this loop is the
// single-GPU replacement of that pipeline's search stage.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <random>

#include "../../include/mivf.h"
#include "index_impl.h"

namespace mivf {
namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
    const float *d_q;
    int nq, w, k;
    int32_t *d_ids;
    float *d_dis;
    double t_submit = 0, t_done = 0;
    int state = 0;  // 0 waiting, 1 in flight, 2 done, 3 failed at launch
    int64_t mem = 0;
};

struct Slot {
    Ws ws;
    std::shared_ptr<void> xws;  // workspace of the list-major kernels
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    DevBuf q, ids, dis, dw, dk;
    int32_t *h_wk = nullptr;  // pinned: w then k of each query
    int *h_flags = nullptr;   // pinned: list-major overflow count, kernel error flag
    std::vector<int64_t> members;
    bool busy = false, lm = false;
    int nb = 0, kmax = 0;
    int64_t mem = 0;
};

}  // namespace
}  // namespace mivf

using namespace mivf;

struct mivf_server {
    mivf_index *ix = nullptr;
    int nslots = 4, max_batch = 2048;
    int64_t budget = 0, in_flight = 0;
    std::vector<Slot> slots;
    std::vector<Req> reqs;
    std::deque<int64_t> waiting;
    int64_t waiting_q = 0;  // queries of the waiting requests
    int64_t batches = 0, lm_batches = 0, reruns = 0;
    ~mivf_server() {
        for (auto &s : slots) {
            if (s.st) (void)hipStreamSynchronize(s.st);
            s.ws.release();
            for (DevBuf *b : {&s.q, &s.ids, &s.dis, &s.dw, &s.dk}) b->release();
            if (s.ev) (void)hipEventDestroy(s.ev);
            if (s.st) (void)hipStreamDestroy(s.st);
            if (s.h_wk) (void)hipHostFree(s.h_wk);
            if (s.h_flags) (void)hipHostFree(s.h_flags);
            s.xws.reset();
        }
    }
};

namespace {

// Device state a request holds while in flight: its share of the coarse
// distance matrix (with the group minima), probes, residual rows, partial
// top-k lists and the slot's query/output rows.
int64_t request_mem(const mivf_index *ix, int nq, int w, int k) {
    const int64_t per_q = int64_t(coarse_ws_bytes(1, ix->nlist)) + int64_t(w) * 8 + 8 + int64_t(ix->d) * 4 +
                          int64_t(k) * 8 * (1 + 8) + int64_t(w) * (int64_t(ix->d) * 4 + int64_t(ix->m) * 8);
    return per_q * nq;
}

// Enqueue the copies of each member request's rows to its outputs.
int copy_out(mivf_server *sv, Slot &s) {
    int off = 0;
    for (int64_t id : s.members) {
        const Req &r = sv->reqs[id];
        MIVF_CHECK_HIP(hipMemcpy2DAsync(r.d_ids, sizeof(int32_t) * r.k, s.ids.as<int32_t>() + size_t(off) * s.kmax,
                                        sizeof(int32_t) * s.kmax, sizeof(int32_t) * r.k, r.nq,
                                        hipMemcpyDeviceToDevice, s.st));
        MIVF_CHECK_HIP(hipMemcpy2DAsync(r.d_dis, sizeof(float) * r.k, s.dis.as<float>() + size_t(off) * s.kmax,
                                        sizeof(float) * s.kmax, sizeof(float) * r.k, r.nq, hipMemcpyDeviceToDevice,
                                        s.st));
        off += r.nq;
    }
    return kOk;
}

// A finished list-major micro-batch: rerun the queries whose candidates
// overflowed with the plain kernels and copy their requests' rows again
// (rare: the candidate capacity is sized for the bound the kernels keep).
int finish_slot(mivf_server *sv, Slot &s) {
    if (!s.lm) return kOk;
    int nov = 0;
    MIVF_TRY(list_major_finish(sv->ix, s.xws.get(), s.q.as<float>(), s.nb, s.h_wk, s.h_wk + s.nb, s.kmax,
                               s.ids.as<int32_t>(), s.dis.as<float>(), s.h_flags, s.st, &nov));
    if (nov == 0) return kOk;
    sv->reruns += nov;
    MIVF_TRY(copy_out(sv, s));
    MIVF_CHECK_HIP(hipStreamSynchronize(s.st));
    return kOk;
}

int retire(mivf_server *sv, int *completed) {
    for (auto &s : sv->slots) {
        if (!s.busy) continue;
        hipError_t e = hipEventQuery(s.ev);
        if (e == hipErrorNotReady) continue;
        int rc = e == hipSuccess ? finish_slot(sv, s) : kHip;
        if (e != hipSuccess) set_error(std::string("server: micro-batch failed: ") + hipGetErrorString(e));
        const double t = now_ms();
        for (int64_t id : s.members) {
            sv->reqs[id].state = rc == kOk ? 2 : 3;
            sv->reqs[id].t_done = t;
            if (completed && rc == kOk) ++*completed;
        }
        s.members.clear();
        s.busy = false;
        sv->in_flight -= s.mem;
        s.mem = 0;
        if (rc != kOk) return rc;
    }
    return kOk;
}

int launch_taken(mivf_server *sv, Slot &s, const std::vector<int64_t> &take, int nb, int wmax, int kmax);

int launch(mivf_server *sv, Slot &s) {
    int nb = 0, wmax = 1, kmax = 1;
    int64_t mem = 0;
    std::vector<int64_t> take;
    while (!sv->waiting.empty()) {
        const Req &r = sv->reqs[sv->waiting.front()];
        const bool first = take.empty();
        if (!first && nb + r.nq > sv->max_batch) break;
        if (!first && sv->in_flight + mem + r.mem > sv->budget) break;
        if (first && sv->in_flight > 0 && sv->in_flight + r.mem > sv->budget) break;  // wait for memory
        take.push_back(sv->waiting.front());
        sv->waiting.pop_front();
        sv->waiting_q -= r.nq;
        nb += r.nq;
        mem += r.mem;
        wmax = std::max(wmax, r.w);
        kmax = std::max(kmax, r.k);
    }
    if (take.empty()) return kOk;
    const int rc = launch_taken(sv, s, take, nb, wmax, kmax);
    if (rc != kOk) {
        // the requests left the queue: mark them failed so that wait and
        // latency report the error instead of waiting for them
        for (int64_t id : take) sv->reqs[id].state = 3;
        s.members.clear();
        return rc;
    }
    s.members = std::move(take);
    s.busy = true;
    s.mem = mem;
    sv->in_flight += mem;
    ++sv->batches;
    if (s.lm) ++sv->lm_batches;
    return kOk;
}

// Size the slot's buffers for the taken requests and enqueue their search.
int launch_taken(mivf_server *sv, Slot &s, const std::vector<int64_t> &take, int nb, int wmax, int kmax) {
    mivf_index *ix = sv->ix;
    const int d = ix->d;
    MIVF_TRY(s.q.ensure(sizeof(float) * size_t(nb) * d));
    MIVF_TRY(s.ids.ensure(sizeof(int32_t) * size_t(nb) * kmax));
    MIVF_TRY(s.dis.ensure(sizeof(float) * size_t(nb) * kmax));
    MIVF_TRY(s.dw.ensure(sizeof(int32_t) * nb));
    MIVF_TRY(s.dk.ensure(sizeof(int32_t) * nb));
    if (!s.h_wk || nb > sv->max_batch) {
        if (s.h_wk) MIVF_CHECK_HIP(hipHostFree(s.h_wk));
        s.h_wk = nullptr;
        MIVF_CHECK_HIP(hipHostMalloc(reinterpret_cast<void **>(&s.h_wk), sizeof(int32_t) * 2 * size_t(std::max(nb, sv->max_batch))));
    }
    int off = 0;
    for (int64_t id : take) {
        Req &r = sv->reqs[id];
        MIVF_CHECK_HIP(hipMemcpyAsync(s.q.as<float>() + size_t(off) * d, r.d_q, sizeof(float) * size_t(r.nq) * d,
                                      hipMemcpyDeviceToDevice, s.st));
        for (int i = 0; i < r.nq; ++i) {
            s.h_wk[off + i] = r.w;
            s.h_wk[nb + off + i] = r.k;
        }
        r.state = 1;
        off += r.nq;
    }
    MIVF_CHECK_HIP(hipMemcpyAsync(s.dw.p, s.h_wk, sizeof(int32_t) * nb, hipMemcpyHostToDevice, s.st));
    MIVF_CHECK_HIP(hipMemcpyAsync(s.dk.p, s.h_wk + nb, sizeof(int32_t) * nb, hipMemcpyHostToDevice, s.st));
    s.nb = nb;
    s.kmax = kmax;
    s.lm = list_major_fits(ix, nb);
    if (s.lm)
        MIVF_TRY(list_major_enqueue(ix, s.xws.get(), s.q.as<float>(), nb, wmax, s.dw.as<int32_t>(),
                                    s.dk.as<int32_t>(), kmax, s.ids.as<int32_t>(), s.dis.as<float>(), s.h_flags,
                                    s.st));
    else
        MIVF_TRY(search_entry(ix, s.ws, s.q.as<float>(), nb, wmax, s.dw.as<int32_t>(), s.dk.as<int32_t>(), kmax,
                              kmax, s.ids.as<int32_t>(), s.dis.as<float>(), s.st));
    s.members = take;
    MIVF_TRY(copy_out(sv, s));
    MIVF_CHECK_HIP(hipEventRecord(s.ev, s.st));
    return kOk;
}

// Whether a free slot should take the waiting requests now: always when no
// micro-batch is in flight; otherwise once serve_min_fill queries wait or the
// oldest request has waited serve_max_wait_us, so that under load the slots
// run fewer, larger micro-batches.
bool ready_to_launch(const mivf_server *sv) {
    const mivf_index *ix = sv->ix;
    if (ix->serve_min_fill <= 0) return true;
    bool any_busy = false;
    for (const auto &s : sv->slots) any_busy = any_busy || s.busy;
    if (!any_busy || sv->waiting_q >= ix->serve_min_fill) return true;
    const double waited_ms = now_ms() - sv->reqs[sv->waiting.front()].t_submit;
    return waited_ms * 1000.0 >= double(ix->serve_max_wait_us);
}

int progress_impl(mivf_server *sv, int *completed) {
    MIVF_TRY(retire(sv, completed));
    for (auto &s : sv->slots) {
        if (sv->waiting.empty() || !ready_to_launch(sv)) break;
        if (!s.busy) MIVF_TRY(launch(sv, s));
    }
    return kOk;
}

}  // namespace

int mivf_server_create(mivf_index *ix, int slots, int max_batch, int64_t mem_budget, mivf_server **out) try {
    MIVF_REQUIRE(ix && out, kInvalid, "server_create: null argument");
    MIVF_REQUIRE(slots >= 1 && slots <= 16, kInvalid, "server_create: slots must be in 1..16");
    MIVF_REQUIRE(max_batch >= 1 && max_batch <= (1 << 16), kInvalid, "server_create: max_batch must be in 1..65536");
    MIVF_TRY(upload_lists(ix));
    auto *sv = new mivf_server();
    sv->ix = ix;
    sv->nslots = slots;
    sv->max_batch = max_batch;
    sv->budget = mem_budget > 0 ? mem_budget : (int64_t(4) << 30);
    sv->slots.resize(slots);
    for (auto &s : sv->slots) {
        if (hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void **>(&s.h_flags), 2 * sizeof(int)) != hipSuccess) {
            delete sv;
            set_error("server_create: stream/event creation failed");
            return kHip;
        }
        // a slot's buffers are used on its stream only: they grow stream-ordered,
        // so a larger batch never waits for the other slots' work
        s.ws.order_on(s.st);
        for (DevBuf *b : {&s.q, &s.ids, &s.dis, &s.dw, &s.dk}) b->ord = s.st;
        s.xws = new_list_major_ws(s.st);
    }
    *out = sv;
    return kOk;
} MIVF_CATCH_ALL

int mivf_server_free(mivf_server *sv) try {
    delete sv;
    return kOk;
} MIVF_CATCH_ALL

int mivf_server_submit(mivf_server *sv, const float *d_q, int nq, int w, int k, int32_t *d_ids, float *d_dis,
                       int64_t *ticket) try {
    MIVF_REQUIRE(sv && ticket, kInvalid, "server_submit: null argument");
    MIVF_REQUIRE(nq >= 1 && d_q && d_ids && d_dis, kInvalid, "server_submit: empty request or null pointer");
    MIVF_REQUIRE(w >= 1 && w <= sv->ix->nlist && w <= 1024 && k >= 1 && k <= 1024, kInvalid,
                 "server_submit: w must be in 1..min(nlist,1024), k in 1..1024");
    Req r;
    r.d_q = d_q;
    r.nq = nq;
    r.w = w;
    r.k = k;
    r.d_ids = d_ids;
    r.d_dis = d_dis;
    r.t_submit = now_ms();
    r.mem = request_mem(sv->ix, nq, w, k);
    // a micro-batch of this server may take the list-major kernels: charge
    // their candidate buffer as well
    if (list_major_fits(sv->ix, sv->max_batch)) r.mem += int64_t(nq) * list_major_query_bytes(sv->ix, k, w);
    *ticket = int64_t(sv->reqs.size());
    sv->reqs.push_back(r);
    sv->waiting.push_back(*ticket);
    sv->waiting_q += nq;
    return progress_impl(sv, nullptr);
} MIVF_CATCH_ALL

int mivf_server_progress(mivf_server *sv, int *completed) try {
    MIVF_REQUIRE(sv, kInvalid, "server_progress: null server");
    int c = 0;
    MIVF_TRY(progress_impl(sv, &c));
    if (completed) *completed = c;
    return kOk;
} MIVF_CATCH_ALL

int mivf_server_wait(mivf_server *sv, int64_t ticket) try {
    MIVF_REQUIRE(sv && ticket >= 0 && ticket < int64_t(sv->reqs.size()), kInvalid, "server_wait: bad ticket");
    for (;;) {
        const int st = sv->reqs[ticket].state;
        if (st == 2) return kOk;
        MIVF_REQUIRE(st != 3, kHip, "server_wait: the request's micro-batch failed to launch");
        MIVF_TRY(progress_impl(sv, nullptr));
    }
} MIVF_CATCH_ALL

int mivf_server_drain(mivf_server *sv) try {
    MIVF_REQUIRE(sv, kInvalid, "server_drain: null server");
    for (;;) {
        MIVF_TRY(progress_impl(sv, nullptr));
        bool busy = !sv->waiting.empty();
        for (auto &s : sv->slots) busy = busy || s.busy;
        if (!busy) return kOk;
    }
} MIVF_CATCH_ALL

int mivf_server_latency(mivf_server *sv, int64_t ticket, double *ms) try {
    MIVF_REQUIRE(sv && ms && ticket >= 0 && ticket < int64_t(sv->reqs.size()), kInvalid, "server_latency: bad ticket");
    const Req &r = sv->reqs[ticket];
    MIVF_REQUIRE(r.state != 3, kHip, "server_latency: the request's micro-batch failed to launch");
    MIVF_REQUIRE(r.state == 2, kInvalid, "server_latency: request not finished");
    *ms = r.t_done - r.t_submit;
    return kOk;
} MIVF_CATCH_ALL

int mivf_server_stats(mivf_server *sv, int64_t *batches, int64_t *list_major_batches, int64_t *rerun_queries) try {
    MIVF_REQUIRE(sv, kInvalid, "server_stats: null server");
    if (batches) *batches = sv->batches;
    if (list_major_batches) *list_major_batches = sv->lm_batches;
    if (rerun_queries) *rerun_queries = sv->reruns;
    return kOk;
} MIVF_CATCH_ALL

namespace {

// Grow every slot's workspaces before the clock starts: one micro-batch per
// slot of the largest w and k, at the largest size the per-query kernels take
// and at max_batch (as far as the pool allows), so that no timed micro-batch
// allocates device memory.
int warm_up(mivf_server *sv, const float *d_pool, int64_t pool_n, int w, int k) {
    const int full = int(std::min<int64_t>(sv->max_batch, pool_n));
    std::vector<int> sizes;
    const int64_t lm_min = sv->ix->list_major_min;
    if (list_major_fits(sv->ix, full) && lm_min > 1) sizes.push_back(int(std::min<int64_t>(full, lm_min - 1)));
    sizes.push_back(full);
    const size_t per_slot = size_t(full) * k;
    DevBuf oi, od;
    MIVF_TRY(oi.ensure(sizeof(int32_t) * per_slot * sv->slots.size()));
    MIVF_TRY(od.ensure(sizeof(float) * per_slot * sv->slots.size()));
    for (int nq : sizes) {
        for (size_t i = 0; i < sv->slots.size(); ++i) {
            if (sv->slots[i].busy) continue;
            Req r;
            r.d_q = d_pool;
            r.nq = nq;
            r.w = w;
            r.k = k;
            r.d_ids = oi.as<int32_t>() + i * per_slot;
            r.d_dis = od.as<float>() + i * per_slot;
            r.t_submit = now_ms();
            r.mem = request_mem(sv->ix, nq, w, k);
            if (list_major_fits(sv->ix, sv->max_batch)) r.mem += int64_t(nq) * list_major_query_bytes(sv->ix, k, w);
            sv->reqs.push_back(r);
            sv->waiting.push_back(int64_t(sv->reqs.size()) - 1);
            sv->waiting_q += nq;
            MIVF_TRY(launch(sv, sv->slots[i]));
        }
        MIVF_TRY(mivf_server_drain(sv));
    }
    MIVF_CHECK_HIP(hipDeviceSynchronize());
    sv->batches = sv->lm_batches = sv->reruns = 0;
    sv->reqs.clear();
    return kOk;
}

}  // namespace

int mivf_server_warm(mivf_server *sv, const float *d_rows, int64_t n_rows, int w, int k) try {
    MIVF_REQUIRE(sv && d_rows && n_rows >= 1, kInvalid, "server_warm: null argument or no rows");
    MIVF_REQUIRE(w >= 1 && w <= sv->ix->nlist && w <= 1024 && k >= 1 && k <= 1024, kInvalid,
                 "server_warm: w must be in 1..min(nlist,1024), k in 1..1024");
    bool busy = !sv->waiting.empty();
    for (const auto &s : sv->slots) busy = busy || s.busy;
    MIVF_REQUIRE(!busy, kInvalid, "server_warm: requests are waiting or in flight");
    return warm_up(sv, d_rows, n_rows, w, k);
} MIVF_CATCH_ALL

// Open-loop load: nreq requests with exponential inter-arrival times at
// `rate` requests/s, nq log-uniform in 1..bmax, w and k drawn from the given
// choices, query rows at random offsets of d_pool.  Latency = completion -
// scheduled arrival (a request is submitted as soon as its time has come).
// The slots' workspaces are grown to their largest micro-batches first.
int mivf_serve_bench(mivf_index *ix, int slots, int max_batch, int64_t mem_budget, const float *d_pool,
                     int64_t pool_n, double rate, int nreq, uint64_t seed, int bmax, const int32_t *w_choices,
                     int nw, const int32_t *k_choices, int nk, double *stats) try {
    MIVF_REQUIRE(ix && d_pool && stats && nreq >= 1 && rate > 0 && bmax >= 1 && pool_n >= bmax && nw >= 1 &&
                     nk >= 1 && w_choices && k_choices,
                 kInvalid, "serve_bench: bad arguments");
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate / 1000.0);  // per ms
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    struct Plan {
        double t;
        int nq, w, k;
        int64_t row;
        int64_t out;
    };
    std::vector<Plan> plan(nreq);
    double t = 0;
    int64_t outs = 0;
    for (auto &p : plan) {
        t += gap(rng);
        p.t = t;
        p.nq = std::max(1, std::min(bmax, int(std::floor(std::exp(u01(rng) * std::log(double(bmax) + 1.0))))));
        p.w = w_choices[rng() % uint64_t(nw)];
        p.k = k_choices[rng() % uint64_t(nk)];
        p.row = int64_t(rng() % uint64_t(pool_n - p.nq + 1));
        p.out = outs;
        outs += int64_t(p.nq) * p.k;
    }
    DevBuf oids, odis;
    MIVF_TRY(oids.ensure(sizeof(int32_t) * size_t(outs)));
    MIVF_TRY(odis.ensure(sizeof(float) * size_t(outs)));
    mivf_server *sv = nullptr;
    MIVF_TRY(mivf_server_create(ix, slots, max_batch, mem_budget, &sv));
    {
        const int rcw = mivf_server_warm(sv, d_pool, pool_n, *std::max_element(w_choices, w_choices + nw),
                                         *std::max_element(k_choices, k_choices + nk));
        if (rcw != kOk) {
            delete sv;
            return rcw;
        }
    }
    std::vector<int64_t> ticket(nreq);
    int next = 0;
    const double t0 = now_ms();
    int rc = kOk;
    while (rc == kOk) {
        const double tn = now_ms() - t0;
        while (next < nreq && plan[next].t <= tn) {
            const Plan &p = plan[next];
            rc = mivf_server_submit(sv, d_pool + p.row * ix->d, p.nq, p.w, p.k, oids.as<int32_t>() + p.out,
                                    odis.as<float>() + p.out, &ticket[next]);
            if (rc != kOk) break;
            sv->reqs[ticket[next]].t_submit = t0 + p.t;  // latency from the scheduled arrival
            ++next;
        }
        if (rc != kOk) break;
        rc = progress_impl(sv, nullptr);
        if (next == nreq) {
            bool busy = !sv->waiting.empty();
            for (auto &s : sv->slots) busy = busy || s.busy;
            if (!busy) break;
        }
    }
    if (rc != kOk) {
        delete sv;
        return rc;
    }
    std::vector<double> lat(nreq);
    int64_t nqt = 0;
    double tend = 0;
    for (int i = 0; i < nreq; ++i) {
        const Req &r = sv->reqs[ticket[i]];
        lat[i] = r.t_done - r.t_submit;
        nqt += r.nq;
        tend = std::max(tend, r.t_done - t0);
    }
    std::vector<double> sorted = lat;
    std::sort(sorted.begin(), sorted.end());
    auto pct = [&](double p) { return sorted[std::min<size_t>(sorted.size() - 1, size_t(p * (sorted.size() - 1) + 0.5))]; };
    double mean = 0;
    for (double v : lat) mean += v;
    mean /= nreq;
    stats[0] = double(nqt) / (tend / 1000.0);             // achieved queries/s
    stats[1] = pct(0.50);                                 // p50 ms
    stats[2] = pct(0.99);                                 // p99 ms
    stats[3] = mean;                                      // mean ms
    stats[4] = double(nqt);                               // queries served
    stats[5] = tend;                                      // makespan ms
    stats[6] = double(sv->batches);                       // micro-batches launched
    stats[7] = double(nqt) / (plan.back().t / 1000.0);    // offered queries/s
    stats[8] = double(sv->lm_batches);                    // of them on the list-major kernels
    stats[9] = double(sv->reruns);                        // queries rerun after an overflow
    delete sv;
    return kOk;
} MIVF_CATCH_ALL
