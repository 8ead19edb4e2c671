// Host layer of libmivf: index lifecycle, GPU-assisted training and encoding,
// search orchestration and the C ABI declared in include/mivf.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/mivf.h"
#include "index_impl.h"

namespace mivf {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

}  // namespace mivf

using namespace mivf;

namespace mivf {

int upload_lists(mivf_index *ix) {
    if (!ix->dirty) return kOk;
    std::vector<int64_t> off(ix->nlist + 1, 0);
    for (int l = 0; l < ix->nlist; ++l) off[l + 1] = off[l] + int64_t(ix->h_ids[l].size());
    const int64_t n = off[ix->nlist];
    MIVF_TRY(ix->off.ensure(sizeof(int64_t) * (ix->nlist + 1)));
    MIVF_TRY(ix->ids.ensure(sizeof(int32_t) * std::max<int64_t>(n, 1)));
    // 64 bytes of tail padding: the scan issues clamped, branch-free code loads
    MIVF_TRY(ix->codes.ensure(size_t(ix->m) * std::max<int64_t>(n, 1) + 64));
    MIVF_CHECK_HIP(hipMemcpy(ix->off.p, off.data(), sizeof(int64_t) * (ix->nlist + 1), hipMemcpyHostToDevice));
    if (n > 0) {
        std::vector<int32_t> ids(n);
        std::vector<uint8_t> codes(size_t(n) * ix->m);
        for (int l = 0; l < ix->nlist; ++l) {
            std::copy(ix->h_ids[l].begin(), ix->h_ids[l].end(), ids.begin() + off[l]);
            std::copy(ix->h_codes[l].begin(), ix->h_codes[l].end(), codes.begin() + off[l] * ix->m);
        }
        MIVF_CHECK_HIP(hipMemcpy(ix->ids.p, ids.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        MIVF_CHECK_HIP(hipMemcpy(ix->codes.p, codes.data(), codes.size(), hipMemcpyHostToDevice));
    }
    // the lists by decreasing length (ties by index): the order in which the
    // list-major kernels number their work items when option item_order = 1
    {
        std::vector<int32_t> order(ix->nlist);
        for (int l = 0; l < ix->nlist; ++l) order[l] = l;
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t a, int32_t b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
        MIVF_TRY(ix->lorder.ensure(sizeof(int32_t) * ix->nlist));
        MIVF_CHECK_HIP(hipMemcpy(ix->lorder.p, order.data(), sizeof(int32_t) * ix->nlist, hipMemcpyHostToDevice));
    }
    ix->h_off = off;
    ix->ntotal = n;
    ix->dirty = false;
    ix->list_gen++;
    ix->rows.release();  // list positions changed: an attached copy is stale
    ix->rows_ok = false;
    ix->rows_u8 = false;
    return kOk;
}

}  // namespace mivf

namespace {

// Run fn(j, nth) for j < nth on nth host threads (at most the hardware's).
template <typename Fn>
void host_parallel(int64_t want, Fn fn) {
    const int hw = int(std::max(1u, std::thread::hardware_concurrency()));
    const int nth = int(std::max<int64_t>(1, std::min<int64_t>(want, hw)));
    std::vector<std::thread> th;
    for (int j = 1; j < nth; ++j) th.emplace_back(fn, j, nth);
    fn(0, nth);
    for (auto &t : th) t.join();
}

// ---------------------------------------------------------------------------
// k-means: GPU assignment with the reference distance arithmetic (a = centroid,
// b = point, strict '<' -> lowest index), host update in double precision
// (sums in point order), deterministic split of the largest cluster for
// empty ones.  Identical, step for step, to the CPU oracle's k-means.
// ---------------------------------------------------------------------------
int kmeans_dev(const float *d_x, const float *h_x, int64_t n, int d, int ldx, int k,
               int niter, uint64_t seed, float *cent) {
    MIVF_REQUIRE(k >= 1 && n >= k && d >= 1, kInvalid, "kmeans: need n >= k >= 1");
    // seeded distinct initial points (partial Fisher-Yates, Lemire reduction)
    std::vector<int64_t> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = i;
    const uint64_t base = synth_stream_base(seed, 0x6b6d65616e73ull);
    for (int64_t i = 0; i < k; ++i) {
        uint64_t r = synth_draw(base, uint64_t(i));
        int64_t j = i + int64_t((unsigned __int128)r * uint64_t(n - i) >> 64);
        std::swap(perm[i], perm[j]);
    }
    for (int c = 0; c < k; ++c)
        for (int t = 0; t < d; ++t) cent[int64_t(c) * d + t] = h_x[perm[c] * ldx + t];

    DevBuf dc, dcn, dxn, dkeys, dasg;
    MIVF_TRY(dc.ensure(sizeof(float) * size_t(k) * d));
    MIVF_TRY(dcn.ensure(sizeof(float) * k));
    MIVF_TRY(dxn.ensure(sizeof(double) * n));
    MIVF_TRY(dkeys.ensure(sizeof(uint64_t) * n));
    MIVF_TRY(dasg.ensure(sizeof(int32_t) * n));
    MIVF_CHECK_HIP(launch_norm_f64(d_x, n, d, ldx, dxn.as<double>(), nullptr));
    std::vector<int32_t> assign(n);
    std::vector<double> sums(size_t(k) * d);
    std::vector<int64_t> cnt(k);
    for (int it = 0; it < niter; ++it) {
        MIVF_CHECK_HIP(hipMemcpy(dc.p, cent, sizeof(float) * size_t(k) * d, hipMemcpyHostToDevice));
        MIVF_CHECK_HIP(launch_norm_f32(dc.as<float>(), k, d, d, dcn.as<float>(), nullptr));
        MIVF_CHECK_HIP(launch_fill_u64(dkeys.as<uint64_t>(), n, kNoKey, nullptr));
        MIVF_CHECK_HIP(launch_dist_argmin(dc.as<float>(), d, dcn.as<float>(), k, d_x, ldx,
                                          dxn.as<double>(), n, d, dkeys.as<uint64_t>(), nullptr));
        MIVF_CHECK_HIP(launch_keys_to_idx(dkeys.as<uint64_t>(), n, dasg.as<int32_t>(), nullptr));
        MIVF_CHECK_HIP(hipMemcpy(assign.data(), dasg.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i)
            MIVF_REQUIRE(assign[i] >= 0 && assign[i] < k, kInvalid, "kmeans: non-finite distances in training data");
        std::fill(sums.begin(), sums.end(), 0.0);
        std::fill(cnt.begin(), cnt.end(), 0);
        // each host thread owns the clusters c with c % nth == j and adds its
        // points in point order: the same sums as one sequential pass
        host_parallel(std::min<int64_t>(k, 16), [&](int j, int nth) {
            for (int64_t i = 0; i < n; ++i) {
                const int c = assign[i];
                if (c % nth != j) continue;
                cnt[c]++;
                double *s = sums.data() + int64_t(c) * d;
                const float *xi = h_x + i * ldx;
                for (int t = 0; t < d; ++t) s[t] += double(xi[t]);
            }
        });
        for (int c = 0; c < k; ++c)
            if (cnt[c] > 0)
                for (int t = 0; t < d; ++t)
                    cent[int64_t(c) * d + t] = float(sums[int64_t(c) * d + t] / double(cnt[c]));
        for (int c = 0; c < k; ++c) {
            if (cnt[c] > 0) continue;
            int big = 0;
            for (int j = 1; j < k; ++j)
                if (cnt[j] > cnt[big]) big = j;
            for (int t = 0; t < d; ++t) {
                double v = double(cent[int64_t(big) * d + t]);
                double delta = (1.0 / 1024.0) * (1.0 + std::fabs(v));
                cent[int64_t(c) * d + t] = float(v + delta);
                cent[int64_t(big) * d + t] = float(v - delta);
            }
            cnt[c] = cnt[big] / 2;
            cnt[big] -= cnt[c];
        }
    }
    MIVF_CHECK_HIP(hipDeviceSynchronize());
    return kOk;
}

int train_impl(const float *d_x, const float *h_x, int64_t n, int d, int nlist, int m, int ks,
               int niter_c, int niter_pq, uint64_t seed, float *coarse_out, float *pq_out) {
    MIVF_REQUIRE(d > 0 && m > 0 && d % m == 0, kInvalid, "train: d must be a multiple of m");
    MIVF_REQUIRE(ks >= 1 && ks <= 256, kInvalid, "train: ks must be in 1..256");
    MIVF_REQUIRE(n >= nlist && n >= ks, kInvalid, "train: need n >= nlist and n >= ks");
    const int ds = d / m;
    MIVF_TRY(kmeans_dev(d_x, h_x, n, d, d, nlist, niter_c, seed, coarse_out));
    // assign training points, residuals
    DevBuf dc, dcn, dxn, dkeys, dasg, dres;
    MIVF_TRY(dc.ensure(sizeof(float) * size_t(nlist) * d));
    MIVF_TRY(dcn.ensure(sizeof(float) * nlist));
    MIVF_TRY(dxn.ensure(sizeof(double) * n));
    MIVF_TRY(dkeys.ensure(sizeof(uint64_t) * n));
    MIVF_TRY(dasg.ensure(sizeof(int32_t) * n));
    MIVF_TRY(dres.ensure(sizeof(float) * size_t(n) * d));
    MIVF_CHECK_HIP(hipMemcpy(dc.p, coarse_out, sizeof(float) * size_t(nlist) * d, hipMemcpyHostToDevice));
    MIVF_CHECK_HIP(launch_norm_f32(dc.as<float>(), nlist, d, d, dcn.as<float>(), nullptr));
    MIVF_CHECK_HIP(launch_norm_f64(d_x, n, d, d, dxn.as<double>(), nullptr));
    MIVF_CHECK_HIP(launch_fill_u64(dkeys.as<uint64_t>(), n, kNoKey, nullptr));
    MIVF_CHECK_HIP(launch_dist_argmin(dc.as<float>(), d, dcn.as<float>(), nlist, d_x, d,
                                      dxn.as<double>(), n, d, dkeys.as<uint64_t>(), nullptr));
    MIVF_CHECK_HIP(launch_keys_to_idx(dkeys.as<uint64_t>(), n, dasg.as<int32_t>(), nullptr));
    MIVF_CHECK_HIP(launch_residual(d_x, n, d, dc.as<float>(), dasg.as<int32_t>(), dres.as<float>(), nullptr));
    dxn.release();
    dkeys.release();
    std::vector<float> hres(size_t(n) * d);
    MIVF_CHECK_HIP(hipMemcpy(hres.data(), dres.p, sizeof(float) * hres.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < m; ++s) {
        MIVF_TRY(kmeans_dev(dres.as<float>() + s * ds, hres.data() + s * ds, n, ds, d, ks, niter_pq,
                            seed + 1 + uint64_t(s), pq_out + int64_t(s) * ks * ds));
    }
    return kOk;
}

// encode n rows of d_x: coarse argmin, residual, per-subquantizer argmin
int encode_impl(mivf_index *ix, const float *d_x, int64_t n, int32_t *d_assign, uint8_t *d_codes,
                hipStream_t st) {
    const int d = ix->d, m = ix->m, ds = ix->ds;
    MIVF_TRY(ix->ws.qn.ensure(sizeof(double) * std::max<int64_t>(n, 1)));
    MIVF_TRY(ix->keys.ensure(sizeof(uint64_t) * std::max<int64_t>(n, 1)));
    MIVF_TRY(ix->res.ensure(sizeof(float) * size_t(std::max<int64_t>(n, 1)) * d));
    MIVF_TRY(ix->resn.ensure(sizeof(double) * std::max<int64_t>(n, 1)));
    uint64_t *keys = ix->keys.as<uint64_t>();
    MIVF_CHECK_HIP(launch_norm_f64(d_x, n, d, d, ix->ws.qn.as<double>(), st));
    MIVF_CHECK_HIP(launch_fill_u64(keys, n, kNoKey, st));
    MIVF_CHECK_HIP(launch_dist_argmin(ix->coarse.as<float>(), d, ix->coarse_n32.as<float>(), ix->nlist,
                                      d_x, d, ix->ws.qn.as<double>(), n, d, keys, st));
    MIVF_CHECK_HIP(launch_keys_to_idx(keys, n, d_assign, st));
    MIVF_CHECK_HIP(launch_residual(d_x, n, d, ix->coarse.as<float>(), d_assign, ix->res.as<float>(), st));
    for (int s = 0; s < m; ++s) {
        const float *rs = ix->res.as<float>() + s * ds;
        MIVF_CHECK_HIP(launch_norm_f64(rs, n, ds, d, ix->resn.as<double>(), st));
        MIVF_CHECK_HIP(launch_fill_u64(keys, n, kNoKey, st));
        MIVF_CHECK_HIP(launch_dist_argmin(ix->pq.as<float>() + int64_t(s) * ix->ks * ds, ds,
                                          ix->pq_n32.as<float>() + s * ix->ks, ix->ks, rs, d,
                                          ix->resn.as<double>(), n, ds, keys, st));
        MIVF_CHECK_HIP(launch_codes_from_keys(keys, n, m, s, d_codes, st));
    }
    return kOk;
}

int add_impl(mivf_index *ix, const float *d_x, int64_t n, int32_t id0) {
    const int64_t CH = 1 << 20;
    DevBuf da, dcodes;
    MIVF_TRY(da.ensure(sizeof(int32_t) * std::min(n, CH)));
    MIVF_TRY(dcodes.ensure(size_t(ix->m) * std::min(n, CH)));
    std::vector<int32_t> a(std::min(n, CH));
    std::vector<uint8_t> c(size_t(ix->m) * std::min(n, CH));
    for (int64_t r0 = 0; r0 < n; r0 += CH) {
        const int64_t nn = std::min(CH, n - r0);
        MIVF_TRY(encode_impl(ix, d_x + r0 * ix->d, nn, da.as<int32_t>(), dcodes.as<uint8_t>(), nullptr));
        MIVF_CHECK_HIP(hipMemcpy(a.data(), da.p, sizeof(int32_t) * nn, hipMemcpyDeviceToHost));
        MIVF_CHECK_HIP(hipMemcpy(c.data(), dcodes.p, size_t(ix->m) * nn, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < nn; ++i) {
            const int l = a[i];
            MIVF_REQUIRE(l >= 0 && l < ix->nlist, kInvalid, "add: non-finite input vector");
            ix->h_ids[l].push_back(id0 + int32_t(r0 + i));
            ix->h_codes[l].insert(ix->h_codes[l].end(), c.begin() + i * ix->m, c.begin() + (i + 1) * ix->m);
        }
    }
    ix->dirty = true;
    return kOk;
}

int set_lists_impl(mivf_index *ix, const int64_t *offsets, const int32_t *ids, const uint8_t *codes8,
                   const int32_t *codes32) {
    MIVF_REQUIRE(offsets && offsets[0] == 0, kInvalid, "set_lists: offsets[0] must be 0");
    for (int l = 0; l < ix->nlist; ++l)
        MIVF_REQUIRE(offsets[l + 1] >= offsets[l], kInvalid, "set_lists: offsets must be non-decreasing");
    const int64_t n = offsets[ix->nlist];
    MIVF_REQUIRE(n < (int64_t(1) << 31), kInvalid, "set_lists: more than 2^31-1 vectors");
    MIVF_REQUIRE(n == 0 || (ids && (codes8 || codes32)), kInvalid, "set_lists: null ids or codes");
    if (codes32)
        for (int64_t i = 0; i < n * ix->m; ++i)
            MIVF_REQUIRE(codes32[i] >= 0 && codes32[i] < ix->ks, kInvalid,
                         "set_lists: code out of range [0, ks) (codes must fit 8 bits)");
    if (codes8 && ix->ks < 256)
        for (int64_t i = 0; i < n * ix->m; ++i)
            MIVF_REQUIRE(codes8[i] < ix->ks, kInvalid, "set_lists: code out of range [0, ks)");
    for (int l = 0; l < ix->nlist; ++l) {
        const int64_t b = offsets[l], e = offsets[l + 1];
        ix->h_ids[l].assign(ids + b, ids + e);
        ix->h_codes[l].resize(size_t(e - b) * ix->m);
        for (int64_t i = 0; i < (e - b) * ix->m; ++i)
            ix->h_codes[l][i] = codes8 ? codes8[b * ix->m + i] : uint8_t(codes32[b * ix->m + i]);
    }
    ix->dirty = true;
    return upload_lists(ix);
}

int check_index(const mivf_index *ix) {
    MIVF_REQUIRE(ix != nullptr, kInvalid, "null index");
    return kOk;
}



int choose_groups(const mivf_index *ix, int nq, int wmax, int *ppg) {
    int p = ix->ppg_opt;
    if (p <= 0) {
        // enough workgroups to fill 256 CUs several times over
        const int target = 2048;
        int groups = nq >= target ? 1 : std::min(wmax, (target + nq - 1) / nq);
        p = (wmax + groups - 1) / groups;
    }
    p = std::max(1, std::min(p, wmax));
    *ppg = p;
    return (wmax + p - 1) / p;
}

// Core search over a chunk of queries already on the device.
//   coarse (q) -> probes -> scan (-> merge)
// resid_mode: d_q holds residuals [nq*w][d] and d_probes_in the list ids.
int search_core(mivf_index *ix, Ws &ws, const float *d_q, int nq, int wmax, const int32_t *d_wq,
                const int32_t *d_kq, int kuni, int kstride, int32_t *d_ids, float *d_dis,
                hipStream_t st, const int32_t *d_probes_in, Timer &tm) {
    const int nlist = ix->nlist, d = ix->d;
    const int32_t *probes = d_probes_in;
    if (!probes) {
        MIVF_TRY(ws.qn.ensure(sizeof(double) * nq));
        MIVF_TRY(ws.dmat.ensure(coarse_ws_bytes(nq, nlist)));
        MIVF_TRY(ws.probes.ensure(sizeof(int32_t) * size_t(nq) * wmax));
        MIVF_TRY(ws.coadis.ensure(sizeof(float) * size_t(nq) * wmax));
        MIVF_TRY(coarse_stage(ix, ws, d_q, nq, wmax, ws.probes.as<int32_t>(), ws.coadis.as<float>(), st, &tm));
        probes = ws.probes.as<int32_t>();
    } else {
        MIVF_TRY(tm.mark());
    }
    int ppg = 0;
    const int ngroups = d_probes_in ? 1 : choose_groups(ix, nq, wmax, &ppg);
    if (d_probes_in) ppg = wmax;
    ScanArgs a{};
    a.queries = d_q;
    a.coarse = ix->coarse.as<float>();
    a.pq = ix->pq.as<float>();
    a.pq_norm = ix->pq_n64.as<double>();
    a.list_off = ix->off.as<int64_t>();
    a.list_ids = ix->ids.as<int32_t>();
    a.codes = ix->codes.as<uint8_t>();
    a.probes = probes;
    a.wq = d_wq;
    a.kq = d_kq;
    a.nq = nq;
    a.w = wmax;
    a.wstride = wmax;
    a.k = kuni;
    a.kstride = kstride;
    a.ks = ix->ks;
    a.nlist = nlist;
    a.ppg = ppg;
    a.ngroups = ngroups;
    a.resid_mode = d_probes_in ? 1 : 0;
    a.err = ix->errflag.as<int>();
    if (!d_probes_in && ix->query_order && nq >= 2048 && query_order_ok(nlist)) {
        MIVF_TRY(ws.qorder.ensure(sizeof(int32_t) * size_t(nq)));
        MIVF_CHECK_HIP(launch_query_order(probes, wmax, nq, nlist, ws.qorder.as<int32_t>(), st));
        a.qorder = ws.qorder.as<int32_t>();
    }
    if (!d_probes_in) {
        MIVF_TRY(ws.rres.ensure(sizeof(float) * size_t(nq) * wmax * d));
        a.rres = ws.rres.as<float>();
    }
    MIVF_TRY(ws.rnd.ensure(sizeof(double) * size_t(nq) * wmax * ix->m));
    a.rnd = ws.rnd.as<double>();
    if (ngroups > 1) {
        MIVF_TRY(ws.partial.ensure(sizeof(uint64_t) * size_t(nq) * ngroups * kstride));
        a.partial = ws.partial.as<uint64_t>();
    } else {
        a.out_ids = d_ids;
        a.out_dis = d_dis;
    }
    // the select stage ends with the scan's order and residual rows
    MIVF_CHECK_HIP(launch_scan_prologue(a, ix->m, st));
    MIVF_TRY(tm.mark());
    MIVF_CHECK_HIP(launch_scan(a, ix->m, st));
    if (!ix->scan_checked) {  // once per index: the kernel's LDS-layout precondition
        int err = 0;
        MIVF_CHECK_HIP(hipStreamSynchronize(st));
        MIVF_CHECK_HIP(hipMemcpy(&err, ix->errflag.p, sizeof(int), hipMemcpyDeviceToHost));
        MIVF_REQUIRE(err == 0, kHip, "scan kernel: dynamic LDS block not at address 0 (results invalid)");
        ix->scan_checked = true;
    }
    MIVF_TRY(tm.mark());
    if (ngroups > 1)
        MIVF_CHECK_HIP(launch_merge(ws.partial.as<uint64_t>(), nq, ngroups, kstride, d_kq, kuni, d_ids,
                                    d_dis, st));
    MIVF_TRY(tm.mark());
    return kOk;
}

int finish_timer(mivf_index *ix, Timer &tm) {
    if (!tm.on) return kOk;
    MIVF_CHECK_HIP(hipEventSynchronize(tm.ev[tm.n - 1]));
    for (int i = 0; i + 1 < tm.n && i < 8; ++i) {
        float ms = 0;
        MIVF_CHECK_HIP(hipEventElapsedTime(&ms, tm.ev[i], tm.ev[i + 1]));
        ix->timings[i] += ms;
    }
    return kOk;
}


}  // namespace

namespace mivf {

// Coarse top-w of nq queries into probes/dis [nq][w] (needs ws.dmat sized
// coarse_ws_bytes(nq, nlist)): query norms, distances (timer mark), selection.
int coarse_stage(mivf_index *ix, Ws &ws, const float *d_q, int nq, int w, int32_t *probes, float *dis, hipStream_t st,
                 Timer *tm) {
    const int d = ix->d, nlist = ix->nlist;
    MIVF_TRY(ws.qn.ensure(sizeof(double) * nq));
    if (ix->coarse_mfma && coarse_mfma_ok(d, nlist, w) && (reinterpret_cast<uintptr_t>(d_q) & 15u) == 0) {
        // group minima and query terms follow the matrix in ws.dmat
        const CoarseWs lay = coarse_ws(nq, nlist);
        unsigned char *wb = ws.dmat.as<unsigned char>();
        float *gmin = reinterpret_cast<float *>(wb + lay.gmin);
        const int64_t ldg = (nlist + 15) / 16;
        // with the group select, the distance kernel splits the queries and
        // the select writes their norms: no separate prep kernel
        const bool fused = coarse_group_select_ok(nlist, w);
        MIVF_TRY(ws.redo.ensure(sizeof(int32_t) * (size_t(nq) + 1)));
        int32_t *redo = ws.redo.as<int32_t>();  // [0] count (cleared by the distance kernel), then the queries
        if (!fused) MIVF_CHECK_HIP(launch_coarse_prep(d_q, nq, ws.qn.as<double>(), wb + lay.qhi, wb + lay.qlo, st));
        MIVF_CHECK_HIP(launch_coarse_approx(ix->coarse_hi.p, ix->coarse_lo.p, ix->coarse_n32.as<float>(), nlist,
                                            fused ? d_q : nullptr, wb + lay.qhi, wb + lay.qlo, ws.qn.as<double>(), nq,
                                            ws.dmat.as<float>(), nlist, gmin, ldg, redo, st));
        if (tm) MIVF_TRY(tm->mark());
        MIVF_CHECK_HIP(launch_coarse_select(ix->coarse.as<float>(), ix->coarse_n32.as<float>(), ix->maxcn, nlist,
                                            d_q, ws.qn.as<double>(), nq, w, ws.dmat.as<float>(), nlist, gmin, ldg,
                                            probes, dis, redo + 1, redo, fused ? ws.qn.as<double>() : nullptr, st));
    } else {
        MIVF_CHECK_HIP(launch_norm_f64(d_q, nq, d, d, ws.qn.as<double>(), st));
        MIVF_CHECK_HIP(launch_dist_store(ix->coarse.as<float>(), d, ix->coarse_n32.as<float>(), nlist, d_q, d,
                                         ws.qn.as<double>(), nq, d, ws.dmat.as<float>(), nlist, st));
        if (tm) MIVF_TRY(tm->mark());
        MIVF_CHECK_HIP(launch_row_topk(ws.dmat.as<float>(), nq, nlist, nlist, w, probes, dis, st));
    }
    return kOk;
}

}  // namespace mivf

namespace mivf {

int search_entry(mivf_index *ix, Ws &ws, const float *d_q, int nq, int wmax, const int32_t *d_wq,
                 const int32_t *d_kq, int kuni, int kstride, int32_t *d_ids, float *d_dis,
                 hipStream_t st) {
    MIVF_TRY(upload_lists(ix));
    Timer tm;
    std::fill(ix->timings, ix->timings + 8, 0.0f);
    // bound the coarse distance matrix and the residual rows of the probes
    // to ~512 MiB each per chunk
    const int64_t per_q = std::max<int64_t>(int64_t(coarse_ws_bytes(1, ix->nlist)), int64_t(wmax) * (4 * ix->d + 8 * ix->m));
    const int64_t maxq = std::max<int64_t>(1, (int64_t(512) << 20) / per_q);
    for (int64_t q0 = 0; q0 < nq; q0 += maxq) {
        const int nc = int(std::min<int64_t>(maxq, nq - q0));
        MIVF_TRY(tm.start(ix->profile, st, ix->ev));
        MIVF_TRY(search_core(ix, ws, d_q + q0 * ix->d, nc, wmax, d_wq ? d_wq + q0 : nullptr,
                             d_kq ? d_kq + q0 : nullptr, kuni, kstride, d_ids + q0 * kstride,
                             d_dis + q0 * kstride, st, nullptr, tm));
        MIVF_TRY(finish_timer(ix, tm));
    }
    return kOk;
}

}  // namespace mivf

// ---------------------------------------------------------------------------
// Search without allocation (mivf_search_noalloc): the workspace of one plain
// search carved out of a fixed device array.  Every GPU the library runs on
// holds its own copy of the array (kNoAllocScratch bytes) until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns on that device, never through the symbol's name.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratch = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_at[kMaxDevices] = {};  // per device: the address of `scratch` there

// The current device's scratch address (looked up once per device).
int scratch_address(void **out) {
    int dev = 0;
    MIVF_CHECK_HIP(hipGetDevice(&dev));
    MIVF_REQUIRE(dev >= 0 && dev < kMaxDevices, kUnsupported, "search_noalloc: device index above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_at[dev]) {
        void *p = nullptr;
        MIVF_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        MIVF_REQUIRE(p != nullptr, kHip, "search_noalloc: no device address for the scratch array");
        g_scratch_at[dev] = p;
    }
    *out = g_scratch_at[dev];
    return kOk;
}

// Lay the buffers search_core asks for (nq queries, w probes, k results)
// out from base, 256-byte aligned, as views in ws (base == nullptr: only
// count); returns the bytes used.
size_t noalloc_layout(const mivf_index *ix, int nq, int w, int k, unsigned char *base, Ws &ws) {
    size_t off = 0;
    auto take = [&](DevBuf &b, size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        if (base) b.view(base + off, bytes);
        off += bytes;
    };
    const size_t q = size_t(nq), qw = q * size_t(w);
    int ppg = 0;
    const int ngroups = choose_groups(ix, nq, w, &ppg);
    take(ws.qn, sizeof(double) * q);
    take(ws.dmat, coarse_ws_bytes(int64_t(q), ix->nlist));
    take(ws.probes, sizeof(int32_t) * qw);
    take(ws.coadis, sizeof(float) * qw);
    take(ws.redo, sizeof(int32_t) * (q + 1));
    take(ws.qorder, sizeof(int32_t) * q);
    take(ws.rres, sizeof(float) * qw * size_t(ix->d));
    take(ws.rnd, sizeof(double) * qw * size_t(ix->m));
    take(ws.partial, ngroups > 1 ? sizeof(uint64_t) * q * size_t(ngroups) * size_t(k) : 0);
    return off;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *mivf_last_error(void) { return last_error(); }
int mivf_version(void) { return 100; }

int mivf_device_count(int *count) try {
    MIVF_REQUIRE(count, kInvalid, "null count");
    MIVF_CHECK_HIP(hipGetDeviceCount(count));
    return kOk;
} MIVF_CATCH_ALL
int mivf_malloc(void **p, size_t bytes) try {
    MIVF_REQUIRE(p, kInvalid, "null pointer");
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return kNoMem;
    }
    return kOk;
} MIVF_CATCH_ALL
int mivf_free(void *p) try {
    MIVF_CHECK_HIP(hipFree(p));
    return kOk;
} MIVF_CATCH_ALL
int mivf_memcpy_htod(void *dst, const void *src, size_t bytes) try {
    MIVF_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return kOk;
} MIVF_CATCH_ALL
int mivf_memcpy_dtoh(void *dst, const void *src, size_t bytes) try {
    MIVF_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return kOk;
} MIVF_CATCH_ALL
int mivf_memcpy_dtod(void *dst, const void *src, size_t bytes) try {
    MIVF_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return kOk;
} MIVF_CATCH_ALL
int mivf_synchronize(void) try {
    MIVF_CHECK_HIP(hipDeviceSynchronize());
    return kOk;
} MIVF_CATCH_ALL

int mivf_kmeans(const float *x, int64_t n, int d, int ldx, int k, int niter, uint64_t seed,
                float *cent_out) try {
    MIVF_REQUIRE(x && cent_out && n > 0 && d > 0 && ldx >= d, kInvalid, "kmeans: bad arguments");
    DevBuf dx;
    MIVF_TRY(dx.ensure(sizeof(float) * size_t(n) * ldx));
    MIVF_CHECK_HIP(hipMemcpy(dx.p, x, sizeof(float) * size_t(n) * ldx, hipMemcpyHostToDevice));
    return kmeans_dev(dx.as<float>(), x, n, d, ldx, k, niter, seed, cent_out);
} MIVF_CATCH_ALL

int mivf_train(const float *x, int64_t n, int d, int nlist, int m, int ks, int niter_coarse,
               int niter_pq, uint64_t seed, float *coarse_out, float *pq_out) try {
    MIVF_REQUIRE(x && coarse_out && pq_out && n > 0 && d > 0 && nlist > 0, kInvalid, "train: bad arguments");
    DevBuf dx;
    MIVF_TRY(dx.ensure(sizeof(float) * size_t(n) * d));
    MIVF_CHECK_HIP(hipMemcpy(dx.p, x, sizeof(float) * size_t(n) * d, hipMemcpyHostToDevice));
    return train_impl(dx.as<float>(), x, n, d, nlist, m, ks, niter_coarse, niter_pq, seed, coarse_out, pq_out);
} MIVF_CATCH_ALL

int mivf_train_device(const float *d_x, int64_t n, int d, int nlist, int m, int ks, int niter_coarse,
                      int niter_pq, uint64_t seed, float *coarse_out, float *pq_out) try {
    MIVF_REQUIRE(d_x && coarse_out && pq_out && n > 0 && d > 0 && nlist > 0, kInvalid, "train: bad arguments");
    std::vector<float> hx(size_t(n) * d);
    MIVF_CHECK_HIP(hipMemcpy(hx.data(), d_x, sizeof(float) * hx.size(), hipMemcpyDeviceToHost));
    return train_impl(d_x, hx.data(), n, d, nlist, m, ks, niter_coarse, niter_pq, seed, coarse_out, pq_out);
} MIVF_CATCH_ALL

int mivf_index_create(int d, int nlist, int m, int ks, const float *coarse, const float *pq,
                      mivf_index **out) try {
    MIVF_REQUIRE(out && coarse && pq, kInvalid, "index_create: null argument");
    MIVF_REQUIRE(nlist >= 1 && m >= 1 && d >= 1 && d % m == 0, kInvalid, "index_create: d must be a multiple of m");
    MIVF_REQUIRE(ks >= 1 && ks <= 256, kInvalid, "index_create: ks must be in 1..256 (8-bit codes)");
    MIVF_REQUIRE(d == 128 && (m == 4 || m == 8 || m == 16 || m == 32), kUnsupported,
                 "index_create: compiled search kernels cover d = 128 with m in {4, 8, 16, 32}");
    for (int64_t i = 0; i < int64_t(nlist) * d; ++i)
        MIVF_REQUIRE(std::isfinite(coarse[i]), kInvalid, "index_create: non-finite coarse centroid");
    for (int64_t i = 0; i < int64_t(ks) * d; ++i)
        MIVF_REQUIRE(std::isfinite(pq[i]), kInvalid, "index_create: non-finite PQ centroid");
    mivf_index *ix = new mivf_index();
    ix->d = d;
    ix->nlist = nlist;
    ix->m = m;
    ix->ks = ks;
    ix->ds = d / m;
    ix->h_ids.resize(nlist);
    ix->h_codes.resize(nlist);
    auto fail = [&](int rc) {
        delete ix;
        return rc;
    };
    int rc;
    if ((rc = ix->coarse.ensure(sizeof(float) * size_t(nlist) * d)) ||
        (rc = ix->coarse_n32.ensure(sizeof(float) * nlist)) ||
        (rc = ix->pq.ensure(sizeof(float) * size_t(ks) * d)) ||
        (rc = ix->pq_n64.ensure(sizeof(double) * size_t(m) * ks)) ||
        (rc = ix->pq_n32.ensure(sizeof(float) * size_t(m) * ks)) ||
        (rc = ix->errflag.ensure(sizeof(int))))
        return fail(rc);
    {
        std::vector<float> pm(m, 0.0f);
        for (int s2 = 0; s2 < m; ++s2)
            for (int c = 0; c < ks; ++c) {
                double nn = 0.0;
                for (int t = 0; t < ix->ds; ++t) {
                    const double v = pq[(int64_t(s2) * ks + c) * ix->ds + t];
                    nn += v * v;
                }
                pm[s2] = std::max(pm[s2], float(std::sqrt(nn) * (1.0 + 1e-6)));
            }
        if ((rc = ix->pmax.ensure(sizeof(float) * m))) return fail(rc);
        if (hipMemcpy(ix->pmax.p, pm.data(), sizeof(float) * m, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("index_create: upload failed");
            return fail(kHip);
        }
    }
    if (hipMemset(ix->errflag.p, 0, sizeof(int)) != hipSuccess) {
        set_error("index_create: hipMemset failed");
        return fail(kHip);
    }
    if (hipMemcpy(ix->coarse.p, coarse, sizeof(float) * size_t(nlist) * d, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ix->pq.p, pq, sizeof(float) * size_t(ks) * d, hipMemcpyHostToDevice) != hipSuccess ||
        launch_norm_f32(ix->coarse.as<float>(), nlist, d, d, ix->coarse_n32.as<float>(), nullptr) != hipSuccess ||
        launch_norm_f64(ix->pq.as<float>(), int64_t(m) * ks, ix->ds, ix->ds, ix->pq_n64.as<double>(), nullptr) != hipSuccess ||
        launch_norm_f32(ix->pq.as<float>(), int64_t(m) * ks, ix->ds, ix->ds, ix->pq_n32.as<float>(), nullptr) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        set_error("index_create: HIP upload failed");
        return fail(kHip);
    }
    {
        std::vector<float> cn(nlist);
        if (hipMemcpy(cn.data(), ix->coarse_n32.p, sizeof(float) * nlist, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("index_create: readback failed");
            return fail(kHip);
        }
        for (float v : cn) ix->maxcn = std::max(ix->maxcn, v);
    }
    if ((rc = ix->coarse_hi.ensure(2 * size_t(nlist) * d)) || (rc = ix->coarse_lo.ensure(2 * size_t(nlist) * d)))
        return fail(rc);
    if (launch_coarse_split(ix->coarse.as<float>(), int64_t(nlist) * d, ix->coarse_hi.p, ix->coarse_lo.p, nullptr) !=
            hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        set_error("index_create: coarse split failed");
        return fail(kHip);
    }
    if (fast_pb_bytes(m) > 0) {  // operands of the reduced-precision tables
        if ((rc = ix->fpb.ensure(fast_pb_bytes(m)))) return fail(rc);
        if (launch_fast_prep(ix->pq.as<float>(), ix->pmax.as<float>(), m, ks, ix->fpb.p, nullptr) != hipSuccess ||
            hipDeviceSynchronize() != hipSuccess) {
            set_error("index_create: reduced-precision operand setup failed");
            return fail(kHip);
        }
    }
    {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&ix->ncu, hipDeviceAttributeMultiprocessorCount, dev);  // else the default below
        if (ix->ncu <= 0) ix->ncu = 256;
    }
    if ((rc = upload_lists(ix))) return fail(rc);
    *out = ix;
    return kOk;
} MIVF_CATCH_ALL

int mivf_index_free(mivf_index *ix) try {
    if (!ix) return kOk;
    (void)hipDeviceSynchronize();  // buffers may still be in use by queued work
    delete ix;
    return kOk;
} MIVF_CATCH_ALL

int mivf_index_info(const mivf_index *ix, int *d, int *nlist, int *m, int *ks, int64_t *ntotal) try {
    MIVF_TRY(check_index(ix));
    if (d) *d = ix->d;
    if (nlist) *nlist = ix->nlist;
    if (m) *m = ix->m;
    if (ks) *ks = ix->ks;
    if (ntotal) {
        int64_t n = 0;
        for (auto &v : ix->h_ids) n += int64_t(v.size());
        *ntotal = n;
    }
    return kOk;
} MIVF_CATCH_ALL

int mivf_index_set_lists(mivf_index *ix, const int64_t *offsets, const int32_t *ids, const uint8_t *codes) try {
    MIVF_TRY(check_index(ix));
    return set_lists_impl(ix, offsets, ids, codes, nullptr);
} MIVF_CATCH_ALL

int mivf_index_set_lists_i32(mivf_index *ix, const int64_t *offsets, const int32_t *ids, const int32_t *codes) try {
    MIVF_TRY(check_index(ix));
    return set_lists_impl(ix, offsets, ids, nullptr, codes);
} MIVF_CATCH_ALL

int mivf_index_get_lists(mivf_index *ix, int64_t *offsets, int32_t *ids, uint8_t *codes) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(offsets, kInvalid, "get_lists: null offsets");
    int64_t o = 0;
    offsets[0] = 0;
    for (int l = 0; l < ix->nlist; ++l) {
        const int64_t len = int64_t(ix->h_ids[l].size());
        if (ids) std::memcpy(ids + o, ix->h_ids[l].data(), sizeof(int32_t) * len);
        if (codes) std::memcpy(codes + o * ix->m, ix->h_codes[l].data(), size_t(ix->m) * len);
        o += len;
        offsets[l + 1] = o;
    }
    return kOk;
} MIVF_CATCH_ALL

int mivf_index_add(mivf_index *ix, const float *x, int64_t n, int32_t id0) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(n >= 0 && (n == 0 || x), kInvalid, "add: bad arguments");
    MIVF_REQUIRE(int64_t(id0) + n <= (int64_t(1) << 31), kInvalid, "add: ids exceed int32");
    if (n == 0) return kOk;
    const int64_t CH = 1 << 20;
    DevBuf dx;
    MIVF_TRY(dx.ensure(sizeof(float) * size_t(std::min(n, CH)) * ix->d));
    for (int64_t r0 = 0; r0 < n; r0 += CH) {
        const int64_t nn = std::min(CH, n - r0);
        MIVF_CHECK_HIP(hipMemcpy(dx.p, x + r0 * ix->d, sizeof(float) * size_t(nn) * ix->d, hipMemcpyHostToDevice));
        MIVF_TRY(add_impl(ix, dx.as<float>(), nn, id0 + int32_t(r0)));
    }
    return kOk;
} MIVF_CATCH_ALL

int mivf_index_add_file(mivf_index *ix, const char *path, int elem_bytes, int64_t row0, int64_t n, int32_t id0,
                        int64_t chunk) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(path && (elem_bytes == 1 || elem_bytes == 4) && row0 >= 0 && n >= 0, kInvalid,
                 "index_add_file: bad arguments");
    MIVF_REQUIRE(int64_t(id0) + n <= int64_t(INT32_MAX) + 1, kInvalid, "index_add_file: ids past 2^31 - 1");
    if (n == 0) return kOk;
    const int64_t ch = std::max<int64_t>(1, std::min<int64_t>(chunk > 0 ? chunk : (int64_t(1) << 20), n));
    std::vector<float> h(size_t(ch) * ix->d);
    DevBuf dx;
    MIVF_TRY(dx.ensure(sizeof(float) * h.size()));
    for (int64_t r = 0; r < n; r += ch) {
        const int64_t nn = std::min(ch, n - r);
        MIVF_TRY(read_vecs_rows<float>(path, elem_bytes, row0 + r, nn, ix->d, h.data()));
        MIVF_CHECK_HIP(hipMemcpy(dx.p, h.data(), sizeof(float) * size_t(nn) * ix->d, hipMemcpyHostToDevice));
        MIVF_TRY(add_impl(ix, dx.as<float>(), nn, id0 + int32_t(r)));
    }
    return kOk;
} MIVF_CATCH_ALL

int mivf_index_add_device(mivf_index *ix, const float *d_x, int64_t n, int32_t id0) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(n >= 0 && (n == 0 || d_x), kInvalid, "add: bad arguments");
    MIVF_REQUIRE(int64_t(id0) + n <= (int64_t(1) << 31), kInvalid, "add: ids exceed int32");
    if (n == 0) return kOk;
    return add_impl(ix, d_x, n, id0);
} MIVF_CATCH_ALL

int mivf_index_encode(mivf_index *ix, const float *d_x, int64_t n, int32_t *d_assign, uint8_t *d_codes,
                      void *stream) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(n >= 0 && (n == 0 || (d_x && d_assign && d_codes)), kInvalid, "encode: bad arguments");
    return encode_impl(ix, d_x, n, d_assign, d_codes, static_cast<hipStream_t>(stream));
} MIVF_CATCH_ALL

int mivf_search(mivf_index *ix, const float *d_q, int nq, int w, int k, int32_t *d_ids, float *d_dis,
                void *stream) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(nq >= 0, kInvalid, "search: nq < 0");
    MIVF_REQUIRE(w >= 1 && w <= ix->nlist, kInvalid, "search: w must be in 1..nlist");
    MIVF_REQUIRE(w <= 1024, kUnsupported, "search: w > 1024");
    MIVF_REQUIRE(k >= 1 && k <= 1024, kInvalid, "search: k must be in 1..1024");
    if (nq == 0) return kOk;
    MIVF_REQUIRE(d_q && d_ids && d_dis, kInvalid, "search: null pointer");
    if (list_major_ok(ix, d_q, nq))
        return search_list_major(ix, d_q, nq, w, k, d_ids, d_dis, static_cast<hipStream_t>(stream));
    return search_entry(ix, ix->ws, d_q, nq, w, nullptr, nullptr, k, k, d_ids, d_dis, static_cast<hipStream_t>(stream));
} MIVF_CATCH_ALL

int mivf_search_noalloc(mivf_index *ix, const float *d_q, int nq, int w, int k, int32_t *d_ids, float *d_dis,
                        void *stream) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(nq >= 0, kInvalid, "search_noalloc: nq < 0");
    MIVF_REQUIRE(w >= 1 && w <= ix->nlist, kInvalid, "search_noalloc: w must be in 1..nlist");
    MIVF_REQUIRE(w <= 1024, kUnsupported, "search_noalloc: w > 1024");
    MIVF_REQUIRE(k >= 1 && k <= 1024, kInvalid, "search_noalloc: k must be in 1..1024");
    if (nq == 0) return kOk;
    MIVF_REQUIRE(d_q && d_ids && d_dis, kInvalid, "search_noalloc: null pointer");
    MIVF_REQUIRE(!ix->dirty, kInvalid,
                 "search_noalloc: the lists changed since they were last put on the device; one call of any "
                 "other search (which may allocate) uploads them");
    Ws ws;
    const size_t need = noalloc_layout(ix, nq, w, k, nullptr, ws);
    MIVF_REQUIRE(need <= kNoAllocScratch, kNoMem,
                 "search_noalloc: " + std::to_string(nq) + " queries at w = " + std::to_string(w) + ", k = " +
                     std::to_string(k) + " need " + std::to_string(need) + " bytes of scratch, above the fixed " +
                     std::to_string(kNoAllocScratch) + "; split the batch");
    void *base = nullptr;
    MIVF_TRY(scratch_address(&base));
    noalloc_layout(ix, nq, w, k, static_cast<unsigned char *>(base), ws);
    hipStream_t st = static_cast<hipStream_t>(stream);
    Timer tm;
    MIVF_TRY(tm.start(false, st, nullptr));
    return search_core(ix, ws, d_q, nq, w, nullptr, nullptr, k, k, d_ids, d_dis, st, nullptr, tm);
} MIVF_CATCH_ALL

int mivf_search_var(mivf_index *ix, const float *d_q, int nq, const int32_t *w, const int32_t *k,
                    int32_t *d_ids, float *d_dis, int kmax, void *stream) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(nq >= 0 && kmax >= 1 && kmax <= 1024, kInvalid, "search_var: kmax must be in 1..1024");
    if (nq == 0) return kOk;
    MIVF_REQUIRE(d_q && w && k && d_ids && d_dis, kInvalid, "search_var: null pointer");
    int wmax = 1;
    for (int i = 0; i < nq; ++i) {
        MIVF_REQUIRE(w[i] >= 1 && w[i] <= ix->nlist && w[i] <= 1024, kInvalid, "search_var: w[i] out of range");
        MIVF_REQUIRE(k[i] >= 1 && k[i] <= kmax, kInvalid, "search_var: k[i] out of 1..kmax");
        wmax = std::max(wmax, int(w[i]));
    }
    MIVF_TRY(ix->ws.wq.ensure(sizeof(int32_t) * nq));
    MIVF_TRY(ix->ws.kq.ensure(sizeof(int32_t) * nq));
    MIVF_CHECK_HIP(hipMemcpy(ix->ws.wq.p, w, sizeof(int32_t) * nq, hipMemcpyHostToDevice));
    MIVF_CHECK_HIP(hipMemcpy(ix->ws.kq.p, k, sizeof(int32_t) * nq, hipMemcpyHostToDevice));
    return search_entry(ix, ix->ws, d_q, nq, wmax, ix->ws.wq.as<int32_t>(), ix->ws.kq.as<int32_t>(), kmax, kmax, d_ids,
                        d_dis, static_cast<hipStream_t>(stream));
} MIVF_CATCH_ALL

int mivf_coarse(mivf_index *ix, const float *d_q, int nq, int w, int32_t *d_idx, float *d_dis, void *stream) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(w >= 1 && w <= ix->nlist && w <= 1024, kInvalid, "coarse: w must be in 1..min(nlist,1024)");
    if (nq <= 0) return kOk;
    MIVF_REQUIRE(d_q && d_idx && d_dis, kInvalid, "coarse: null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t maxq = std::max<int64_t>(1, (int64_t(512) << 20) / int64_t(coarse_ws_bytes(1, ix->nlist)));
    for (int64_t q0 = 0; q0 < nq; q0 += maxq) {
        const int nc = int(std::min<int64_t>(maxq, nq - q0));
        MIVF_TRY(ix->ws.qn.ensure(sizeof(double) * nc));
        MIVF_TRY(ix->ws.dmat.ensure(coarse_ws_bytes(nc, ix->nlist)));
        MIVF_TRY(coarse_stage(ix, ix->ws, d_q + q0 * ix->d, nc, w, d_idx + q0 * w, d_dis + q0 * w, st, nullptr));
    }
    return kOk;
} MIVF_CATCH_ALL

int mivf_search_residual(mivf_index *ix, const float *d_residual, const int32_t *d_list_of, int nq, int w,
                         int k, int32_t *d_ids, float *d_dis, void *stream) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(w >= 1 && k >= 1 && k <= 1024, kInvalid, "search_residual: bad w or k");
    if (nq <= 0) return kOk;
    MIVF_REQUIRE(d_residual && d_list_of && d_ids && d_dis, kInvalid, "search_residual: null pointer");
    MIVF_TRY(upload_lists(ix));
    Timer tm;
    MIVF_TRY(tm.start(false, static_cast<hipStream_t>(stream), nullptr));
    return search_core(ix, ix->ws, d_residual, nq, w, nullptr, nullptr, k, k, d_ids, d_dis,
                       static_cast<hipStream_t>(stream), d_list_of, tm);
} MIVF_CATCH_ALL

int mivf_adc(mivf_index *ix, const float *d_residual, const uint8_t *d_codes, int64_t n, float *d_dis,
             void *stream) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(n >= 0, kInvalid, "adc: negative code count");
    if (n == 0) return kOk;
    MIVF_REQUIRE(d_residual && d_codes && d_dis, kInvalid, "adc: null pointer");
    MIVF_CHECK_HIP(launch_adc_codes(d_residual, ix->m, ix->ks, ix->ds, ix->pq.as<float>(), ix->pq_n64.as<double>(),
                                    d_codes, n, d_dis, static_cast<hipStream_t>(stream)));
    return kOk;
} MIVF_CATCH_ALL

int mivf_kmin(const float *d_dis, const int32_t *d_ids, int64_t n, int k, int32_t *d_out_ids, float *d_out_dis,
              void *stream) try {
    MIVF_REQUIRE(k >= 1 && k <= 1024, kInvalid, "kmin: k must be in 1..1024");
    MIVF_REQUIRE(n >= 0 && n <= INT32_MAX, kInvalid, "kmin: n must be in 0..2^31-1");
    MIVF_REQUIRE(d_out_ids && d_out_dis && (n == 0 || (d_dis && d_ids)), kInvalid, "kmin: null pointer");
    MIVF_CHECK_HIP(launch_kmin(d_dis, d_ids, n, k, d_out_ids, d_out_dis, static_cast<hipStream_t>(stream)));
    return kOk;
} MIVF_CATCH_ALL

int mivf_search_rerank(mivf_index *ix, const float *d_q, int nq, int w, int kl, int k, const float *d_base,
                       int64_t nbase, int32_t *d_ids, float *d_dis, void *stream) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(w >= 1 && w <= ix->nlist && w <= 1024, kInvalid, "rerank: w must be in 1..min(nlist,1024)");
    MIVF_REQUIRE(kl >= 1 && kl <= 1024 && k >= 1 && k <= 1024, kInvalid, "rerank: kl and k must be in 1..1024");
    if (nq <= 0) return kOk;
    MIVF_REQUIRE(d_q && d_base && d_ids && d_dis && nbase >= 0, kInvalid, "rerank: null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    MIVF_TRY(ix->ws.tmp_ids.ensure(sizeof(int32_t) * size_t(nq) * kl));
    MIVF_TRY(ix->ws.tmp_dis.ensure(sizeof(float) * size_t(nq) * kl));
    MIVF_TRY(search_entry(ix, ix->ws, d_q, nq, w, nullptr, nullptr, kl, kl, ix->ws.tmp_ids.as<int32_t>(),
                          ix->ws.tmp_dis.as<float>(), st));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ix->profile) {
        MIVF_CHECK_HIP(hipEventCreate(&e0));
        MIVF_CHECK_HIP(hipEventCreate(&e1));
        MIVF_CHECK_HIP(hipEventRecord(e0, st));
    }
    MIVF_CHECK_HIP(launch_rerank(d_base, nbase, d_q, nq, ix->d, ix->ws.tmp_ids.as<int32_t>(), kl, k, d_ids, d_dis, st));
    if (ix->profile) {
        MIVF_CHECK_HIP(hipEventRecord(e1, st));
        MIVF_CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0;
        MIVF_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        ix->timings[4] = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return kOk;
} MIVF_CATCH_ALL

int mivf_search_fast(mivf_index *ix, const float *d_q, int nq, int w, int kl, int k, const float *d_base,
                     int64_t nbase, int32_t *d_ids, float *d_dis, void *stream) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(ix->m == 8 || ix->m == 16, kUnsupported, "search_fast: compiled for m in {8, 16}");
    MIVF_REQUIRE(ix->nlist <= kMaxPairLists, kUnsupported, "search_fast: nlist above 2^20");
    MIVF_REQUIRE(w >= 1 && w <= ix->nlist && w <= 1024, kInvalid, "search_fast: w must be in 1..min(nlist,1024)");
    MIVF_REQUIRE(kl >= 1 && kl <= 1024 && k >= 1 && k <= kl, kInvalid, "search_fast: need 1 <= k <= kl <= 1024");
    if (nq <= 0) return kOk;
    MIVF_REQUIRE(d_q && d_ids && d_dis && nbase >= 0, kInvalid, "search_fast: null pointer");
    MIVF_REQUIRE((reinterpret_cast<uintptr_t>(d_q) & 15u) == 0, kInvalid, "search_fast: d_q must be 16-byte aligned");
    MIVF_TRY(upload_lists(ix));
    MIVF_REQUIRE(d_base || ix->rows_ok, kInvalid,
                 "search_fast: d_base is NULL and no base is attached (mivf_index_attach_base)");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nlist = ix->nlist, d = ix->d;
    const int capq = std::max(ix->fast_cap, 64 * kl);
    MIVF_REQUIRE(double(nq) * nlist * 4.0 <= double(1u << 31), kInvalid,
                 "search_fast: nq * nlist too large for one batch (split the queries)");
    Timer tm;
    std::fill(ix->timings, ix->timings + 8, 0.0f);
    MIVF_TRY(tm.start(ix->profile, st, ix->ev));
    // 1. coarse top-w
    MIVF_TRY(ix->ws.qn.ensure(sizeof(double) * nq));
    MIVF_TRY(ix->ws.dmat.ensure(coarse_ws_bytes(nq, nlist)));
    MIVF_TRY(ix->ws.probes.ensure(sizeof(int32_t) * size_t(nq) * w));
    MIVF_TRY(ix->ws.coadis.ensure(sizeof(float) * size_t(nq) * w));
    const int64_t np = int64_t(nq) * w;
    MIVF_TRY(ix->f_cnt.ensure(sizeof(int) * (nlist + 1)));
    MIVF_TRY(ix->f_poff.ensure(sizeof(int) * (nlist + 1)));
    MIVF_TRY(ix->f_ioff.ensure(sizeof(int) * (nlist + 1)));
    MIVF_TRY(ix->f_fill.ensure(sizeof(int) * (nlist + 1)));
    MIVF_TRY(ix->f_pairq.ensure(sizeof(int32_t) * std::max<int64_t>(np, 1)));
    MIVF_TRY(ix->f_items.ensure(sizeof(FastItem) * std::max<int64_t>(np, 1)));
    MIVF_TRY(ix->f_ctrl.ensure(sizeof(int) * 4));
    MIVF_TRY(ix->f_cand.ensure(sizeof(uint64_t) * size_t(nq) * capq));
    MIVF_TRY(ix->f_ccnt.ensure(sizeof(int) * nq));
    MIVF_TRY(ix->f_rthr.ensure(sizeof(float) * nq));
    MIVF_TRY(ix->f_rhist.ensure(sizeof(int) * size_t(nq) * fast_hist_bins()));
    MIVF_TRY(ix->f_tau.ensure(sizeof(float) * nq));
    int *ctrl = ix->f_ctrl.as<int>();  // [0] total items, [1] next item
    MIVF_TRY(coarse_stage(ix, ix->ws, d_q, nq, w, ix->ws.probes.as<int32_t>(), ix->ws.coadis.as<float>(), st, nullptr));
    MIVF_TRY(tm.mark());
    // 2. (list, query) grouping and the buffer clears on the second stream,
    //    overlapped with 3. the seed thresholds on the caller's stream
    if (!ix->aux) {
        MIVF_CHECK_HIP(hipStreamCreateWithFlags(&ix->aux, hipStreamNonBlocking));
        for (auto &e : ix->aux_ev) MIVF_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    hipStream_t sa = ix->aux;
    MIVF_CHECK_HIP(hipEventRecord(ix->aux_ev[0], st));
    MIVF_CHECK_HIP(hipStreamWaitEvent(sa, ix->aux_ev[0], 0));
    {
        // candidate counts, the work counter, the running bounds (+inf) and
        // their histograms start cleared, off the caller's stream: the fork
        // follows every earlier search on the index, and the scan waits for
        // the second stream
        FillSet f{};
        f.p[0] = ix->f_ccnt.as<int>(), f.n[0] = nq, f.v[0] = 0;
        f.p[1] = ctrl + 1, f.n[1] = 1, f.v[1] = 0;
        f.p[2] = ix->f_rthr.as<int>(), f.n[2] = nq, f.v[2] = 0x7f800000;
        f.p[3] = ix->f_rhist.as<int>(), f.n[3] = int64_t(nq) * fast_hist_bins(), f.v[3] = 0;
        f.count = 4;
        MIVF_CHECK_HIP(launch_fill_multi(f, sa));
    }
    // work items of each query's fast_front nearest lists come first, so that
    // their candidates set the queries' running bounds before the other
    // lists are scanned (the order changes no result)
    const int front = (ix->fast_running && ix->fast_front > 0 && ix->fast_front < w) ? ix->fast_front : 0;
    const int32_t *lorder = ix->item_order ? ix->lorder.as<int32_t>() : nullptr;
    MIVF_CHECK_HIP(launch_pairs(ix->ws.probes.as<int32_t>(), nq, w, 0, front > 0 ? front : w, ix->off.as<int64_t>(),
                                nullptr, nlist, fast_group_size(), ix->f_cnt.as<int>(), ix->f_poff.as<int>(),
                                ix->f_ioff.as<int>(), ctrl, nullptr, lorder, ix->f_fill.as<int>(),
                                ix->f_pairq.as<int32_t>(), ix->f_items.as<FastItem>(), sa));
    if (front > 0)
        MIVF_CHECK_HIP(launch_pairs(ix->ws.probes.as<int32_t>(), nq, w, front, w, ix->off.as<int64_t>(), nullptr,
                                    nlist, fast_group_size(), ix->f_cnt.as<int>(), ix->f_poff.as<int>(),
                                    ix->f_ioff.as<int>(), ctrl, ctrl, lorder, ix->f_fill.as<int>(),
                                    ix->f_pairq.as<int32_t>(), ix->f_items.as<FastItem>(), sa));
    MIVF_CHECK_HIP(hipEventRecord(ix->aux_ev[1], sa));
    // seed threshold from each query's nearest lists (at least fast_seed,
    // then as many as it takes to hold kl vectors)
    {
        SeedArgs a{};
        a.queries = d_q;
        a.coarse = ix->coarse.as<float>();
        a.pq = ix->pq.as<float>();
        a.pnorm = ix->pq_n32.as<float>();
        a.pmax = ix->pmax.as<float>();
        a.list_off = ix->off.as<int64_t>();
        a.codes = ix->codes.as<uint8_t>();
        a.probes = ix->ws.probes.as<int32_t>();
        a.min_seed = std::min(ix->fast_seed, w);
        a.nq = nq;
        a.w = w;
        a.nlist = nlist;
        a.ks = ix->ks;
        a.kl = kl;
        a.cap = ix->fast_seed_cap > 0 ? std::max(ix->fast_seed_cap, kl) : 0;
        a.tau = ix->f_tau.as<float>();
        MIVF_CHECK_HIP(launch_seed_tau(a, ix->m, st));
    }
    MIVF_TRY(tm.mark());
    MIVF_CHECK_HIP(hipStreamWaitEvent(st, ix->aux_ev[1], 0));
    MIVF_TRY(tm.mark());
    // 4. persistent list-major scan
    {
        FastArgs a{};
        a.queries = d_q;
        a.coarse = ix->coarse.as<float>();
        a.pb = ix->fpb.p;
        a.pmax = ix->pmax.as<float>();
        a.codes = ix->codes.as<uint8_t>();
        a.items = ix->f_items.as<FastItem>();
        a.total_items = ctrl;
        a.next_item = ctrl + 1;
        a.cand = ix->f_cand.as<uint64_t>();
        a.cand_cnt = ix->f_ccnt.as<int>();
        a.capq = capq;
        a.ks = ix->ks;
        a.err = ix->errflag.as<int>();
        a.slack = float(ix->fast_slack);
        a.tau = ix->f_tau.as<float>();
        a.runthr = ix->fast_running ? ix->f_rthr.as<float>() : nullptr;
        a.rhist = ix->fast_running ? ix->f_rhist.as<int>() : nullptr;
        a.kl = kl;
        a.pq = ix->pq.as<float>();
        a.pq_norm = ix->pq_n64.as<double>();
        int dev = 0, ncu = 256;
        MIVF_CHECK_HIP(hipGetDevice(&dev));
        MIVF_CHECK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        if (ix->fast_tables == 0) {
            MIVF_CHECK_HIP(launch_fast_scan(a, ix->m, 2 * ncu, st));
        } else {
            MIVF_REQUIRE(fast_hist_bins() == 64, kHip, "search_fast: histogram stride differs between the scans");
            MIVF_CHECK_HIP(launch_fast_var_scan(a, ix->m, ix->fast_tables, 2 * ncu, st));
        }
    }
    MIVF_TRY(tm.mark());
    // 5. shortlist and exact rerank (the coarse stage left the queries' f64
    //    norms in ws.qn); shortlists of up to 256 in one kernel
    if (kl <= 256) {
        const void *rows = d_base ? static_cast<const void *>(d_base) : ix->rows.p;
        MIVF_CHECK_HIP(launch_fast_select_rerank(kl, ix->f_cand.as<uint64_t>(), ix->f_ccnt.as<int>(), capq, nq, rows,
                                                 !d_base && ix->rows_u8, !d_base, d_base ? nbase : ix->ntotal,
                                                 ix->ids.as<int32_t>(), d_q, d, ix->ws.qn.as<double>(), k, d_ids,
                                                 d_dis, st));
        MIVF_TRY(tm.mark());
        MIVF_TRY(tm.mark());  // (no separate rerank stage)
    } else {
        MIVF_TRY(ix->f_short.ensure(sizeof(int32_t) * size_t(nq) * kl));
        MIVF_CHECK_HIP(launch_fast_select(kl, ix->f_cand.as<uint64_t>(), ix->f_ccnt.as<int>(), capq, nq,
                                          ix->f_short.as<int32_t>(), st));
        MIVF_TRY(tm.mark());
        if (d_base)
            MIVF_CHECK_HIP(launch_rerank_pos(d_base, nbase, false, ix->ids.as<int32_t>(), d_q, nq, d,
                                             ix->ws.qn.as<double>(), ix->f_short.as<int32_t>(), kl, k, d_ids, d_dis, st));
        else if (ix->rows_u8)
            MIVF_CHECK_HIP(launch_rerank_pos_u8(ix->rows.as<uint8_t>(), ix->ntotal, ix->ids.as<int32_t>(), d_q, nq, d,
                                                ix->ws.qn.as<double>(), ix->f_short.as<int32_t>(), kl, k, d_ids, d_dis,
                                                st));
        else
            MIVF_CHECK_HIP(launch_rerank_pos(ix->rows.as<float>(), ix->ntotal, true, ix->ids.as<int32_t>(), d_q, nq, d,
                                             ix->ws.qn.as<double>(), ix->f_short.as<int32_t>(), kl, k, d_ids, d_dis, st));
        MIVF_TRY(tm.mark());
    }
    MIVF_TRY(finish_timer(ix, tm));
    if (!ix->scan_checked || ix->profile) {
        int err = 0;
        MIVF_CHECK_HIP(hipStreamSynchronize(st));
        MIVF_CHECK_HIP(hipMemcpy(&err, ix->errflag.p, sizeof(int), hipMemcpyDeviceToHost));
        MIVF_REQUIRE(err == 0, kHip, "kernel precondition failed: dynamic LDS block not at address 0");
        ix->scan_checked = true;
        std::vector<int> cc(nq);
        MIVF_CHECK_HIP(hipMemcpy(cc.data(), ix->f_ccnt.p, sizeof(int) * nq, hipMemcpyDeviceToHost));
        int64_t tot = 0, over = 0;
        for (int v : cc) {
            tot += v;
            over += v > capq;
        }
        ix->f_last_cands = tot;
        ix->f_last_overflow = over;
    }
    return kOk;
} MIVF_CATCH_ALL

int mivf_index_attach_base(mivf_index *ix, const float *d_base, int64_t nbase) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(d_base && nbase >= 0, kInvalid, "attach_base: null base");
    MIVF_TRY(upload_lists(ix));
    int64_t maxid = -1;
    for (const auto &v : ix->h_ids)
        for (int32_t id : v) maxid = std::max<int64_t>(maxid, id);
    MIVF_REQUIRE(maxid < nbase, kInvalid, "attach_base: the lists hold ids >= nbase");
    ix->rows_ok = false;
    ix->rows_u8 = false;
    const size_t nrow = size_t(std::max<int64_t>(ix->ntotal, 1));
    if (ix->base_bytes) {
        // byte copy: exact only when every value is an integer in 0..255
        MIVF_TRY(ix->rows.ensure(nrow * ix->d));
        MIVF_CHECK_HIP(hipMemset(ix->errflag.p, 0, sizeof(int)));
        MIVF_CHECK_HIP(launch_gather_rows_u8(d_base, ix->d, ix->ids.as<int32_t>(), ix->ntotal, ix->rows.as<uint8_t>(),
                                             ix->errflag.as<int>(), nullptr));
        int bad = 0;
        MIVF_CHECK_HIP(hipDeviceSynchronize());
        MIVF_CHECK_HIP(hipMemcpy(&bad, ix->errflag.p, sizeof(int), hipMemcpyDeviceToHost));
        MIVF_CHECK_HIP(hipMemset(ix->errflag.p, 0, sizeof(int)));  // the flag also reports kernel preconditions
        MIVF_REQUIRE(bad == 0, kInvalid,
                     "attach_base: option base_bytes needs every base value to be an integer in 0..255");
        ix->rows_u8 = true;
    } else {
        MIVF_TRY(ix->rows.ensure(sizeof(float) * nrow * ix->d));
        MIVF_CHECK_HIP(launch_gather_rows(d_base, ix->d, ix->ids.as<int32_t>(), ix->ntotal, ix->rows.as<float>(),
                                          nullptr));
        MIVF_CHECK_HIP(hipDeviceSynchronize());
    }
    ix->rows_ok = true;
    return kOk;
} MIVF_CATCH_ALL

int mivf_fast_tables(mivf_index *ix, const float *d_q, int n, int list, uint8_t *d_t8, float *d_delta) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(ix->m == 8 || ix->m == 16, kUnsupported, "fast_tables: compiled for m in {8, 16}");
    MIVF_REQUIRE(n >= 1 && n <= fast_group_size(), kInvalid, "fast_tables: need 1 <= n <= 16 queries");
    MIVF_REQUIRE(list >= 0 && list < ix->nlist, kInvalid, "fast_tables: list out of range");
    MIVF_REQUIRE(d_q && d_t8 && d_delta, kInvalid, "fast_tables: null pointer");
    FastArgs a{};
    a.queries = d_q;
    a.coarse = ix->coarse.as<float>();
    a.pb = ix->fpb.p;
    a.pmax = ix->pmax.as<float>();
    a.ks = ix->ks;
    a.slack = float(ix->fast_slack);
    MIVF_CHECK_HIP(launch_fast_tables(a, ix->m, list, n, d_t8, d_delta, nullptr));
    MIVF_CHECK_HIP(hipDeviceSynchronize());
    return kOk;
} MIVF_CATCH_ALL

int mivf_fast_var_adc(mivf_index *ix, const float *d_q, int n, int list, int tables, float *d_out) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(d_q && d_out && n >= 1 && n <= 16 && list >= 0 && list < ix->nlist, kInvalid,
                 "fast_var_adc: bad arguments");
    MIVF_REQUIRE(tables == 1 || tables == 2, kInvalid, "fast_var_adc: tables must be 1 or 2");
    MIVF_REQUIRE(ix->m == 8 || ix->m == 16, kUnsupported, "fast_var_adc: compiled for m in {8, 16}");
    MIVF_TRY(upload_lists(ix));
    FastArgs a{};
    a.queries = d_q;
    a.coarse = ix->coarse.as<float>();
    a.pmax = ix->pmax.as<float>();
    a.codes = ix->codes.as<uint8_t>();
    a.ks = ix->ks;
    a.err = ix->errflag.as<int>();
    a.pq = ix->pq.as<float>();
    a.pq_norm = ix->pq_n64.as<double>();
    const int64_t beg = ix->h_off[list];
    const int len = int(ix->h_off[list + 1] - beg);
    MIVF_CHECK_HIP(launch_fast_var_adc(a, ix->m, tables, list, n, beg, len, d_out, nullptr));
    MIVF_CHECK_HIP(hipDeviceSynchronize());
    return kOk;
} MIVF_CATCH_ALL

int mivf_fast_seed_tau(mivf_index *ix, const float *d_q, int nq, int w, int kl, float *d_tau) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(ix->m == 8 || ix->m == 16, kUnsupported, "fast_seed_tau: compiled for m in {8, 16}");
    MIVF_REQUIRE(w >= 1 && w <= ix->nlist && w <= 1024 && kl >= 1, kInvalid, "fast_seed_tau: bad w or kl");
    MIVF_REQUIRE(d_q && d_tau && (reinterpret_cast<uintptr_t>(d_q) & 15u) == 0, kInvalid,
                 "fast_seed_tau: null or unaligned pointer");
    if (nq <= 0) return kOk;
    MIVF_TRY(upload_lists(ix));
    MIVF_TRY(ix->ws.dmat.ensure(coarse_ws_bytes(nq, ix->nlist)));
    MIVF_TRY(ix->ws.probes.ensure(sizeof(int32_t) * size_t(nq) * w));
    MIVF_TRY(ix->ws.coadis.ensure(sizeof(float) * size_t(nq) * w));
    MIVF_TRY(coarse_stage(ix, ix->ws, d_q, nq, w, ix->ws.probes.as<int32_t>(), ix->ws.coadis.as<float>(), nullptr, nullptr));
    SeedArgs a{};
    a.queries = d_q;
    a.coarse = ix->coarse.as<float>();
    a.pq = ix->pq.as<float>();
    a.pnorm = ix->pq_n32.as<float>();
    a.pmax = ix->pmax.as<float>();
    a.list_off = ix->off.as<int64_t>();
    a.codes = ix->codes.as<uint8_t>();
    a.probes = ix->ws.probes.as<int32_t>();
    a.min_seed = std::min(ix->fast_seed, w);
    a.nq = nq;
    a.w = w;
    a.nlist = ix->nlist;
    a.ks = ix->ks;
    a.kl = kl;
    a.cap = ix->fast_seed_cap > 0 ? std::max(ix->fast_seed_cap, kl) : 0;
    a.tau = d_tau;
    MIVF_CHECK_HIP(launch_seed_tau(a, ix->m, nullptr));
    MIVF_CHECK_HIP(hipDeviceSynchronize());
    return kOk;
} MIVF_CATCH_ALL

int mivf_fast_stats(mivf_index *ix, int64_t *candidates, int64_t *overflowed_queries) try {
    MIVF_TRY(check_index(ix));
    if (candidates) *candidates = ix->f_last_cands;
    if (overflowed_queries) *overflowed_queries = ix->f_last_overflow;
    return kOk;
} MIVF_CATCH_ALL

int mivf_fast_select(const uint64_t *d_cand, const int32_t *d_cnt, int capq, int nq, int kl, int32_t *d_short) try {
    MIVF_REQUIRE(capq >= 1 && nq >= 0 && kl >= 1, kInvalid, "fast_select: bad sizes");
    MIVF_REQUIRE(d_cand && d_cnt && d_short, kInvalid, "fast_select: null pointer");
    if (nq == 0) return kOk;
    MIVF_CHECK_HIP(launch_fast_select(kl, d_cand, d_cnt, capq, nq, d_short, nullptr));
    MIVF_CHECK_HIP(hipDeviceSynchronize());
    return kOk;
} MIVF_CATCH_ALL

int mivf_set_option(mivf_index *ix, const char *name, int64_t value) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(name, kInvalid, "set_option: null name");
    const std::string n(name);
    if (n == "probes_per_group") {
        MIVF_REQUIRE(value >= 0 && value <= 1024, kInvalid, "probes_per_group must be in 0..1024");
        ix->ppg_opt = int(value);
    } else if (n == "fast_seed_slack") {
        MIVF_REQUIRE(value >= 0 && value <= 64, kInvalid, "fast_seed_slack must be in 0..64");
        ix->fast_slack = int(value);
    } else if (n == "fast_seed_cap") {
        MIVF_REQUIRE(value == 0 || (value >= 1 && value <= 8192), kInvalid, "fast_seed_cap must be 0 or in 1..8192");
        ix->fast_seed_cap = int(value);
    } else if (n == "exact_seed_cap") {
        MIVF_REQUIRE(value == 0 || (value >= 1 && value <= 8192), kInvalid, "exact_seed_cap must be 0 or in 1..8192");
        ix->exact_seed_cap = int(value);
    } else if (n == "fast_running_bound") {
        ix->fast_running = value != 0;
    } else if (n == "query_order") {
        ix->query_order = value != 0;
    } else if (n == "profile") {
        ix->profile = value != 0;
    } else if (n == "fast_seed_lists") {
        MIVF_REQUIRE(value >= 1 && value <= 1024, kInvalid, "fast_seed_lists must be in 1..1024");
        ix->fast_seed = int(value);
    } else if (n == "batched_capacity") {
        MIVF_REQUIRE(value >= 64 && value <= (1 << 20), kInvalid, "batched_capacity must be in 64..2^20");
        ix->batched_cap = int(value);
    } else if (n == "batched_capacity_per_probe") {
        MIVF_REQUIRE(value >= 0 && value <= 65536, kInvalid, "batched_capacity_per_probe must be in 0..65536");
        ix->batched_cap_pp = int(value);
    } else if (n == "batched_group") {
        MIVF_REQUIRE(exact_lds_bytes(ix->m, int(value)) > 0, kInvalid, "batched_group: not a compiled shape");
        ix->batched_g = int(value);
    } else if (n == "batched_threads") {
        MIVF_REQUIRE(value == 256 || value == 512, kInvalid, "batched_threads must be 256 or 512");
        ix->batched_nt = int(value);
    } else if (n == "fast_tables") {
        MIVF_REQUIRE(value >= 0 && value <= 2, kInvalid, "fast_tables must be 0, 1 or 2");
        ix->fast_tables = int(value);
    } else if (n == "exact_running_bound") {
        MIVF_REQUIRE(value >= 0 && value <= 2, kInvalid, "exact_running_bound must be 0, 1 or 2 (automatic)");
        ix->exact_running = int(value);
    } else if (n == "item_order") {
        MIVF_REQUIRE(value == 0 || value == 1, kInvalid, "item_order must be 0 (list order) or 1 (longest lists first)");
        ix->item_order = int(value);
    } else if (n == "fast_front_ranks") {
        MIVF_REQUIRE(value >= 0 && value <= 1024, kInvalid, "fast_front_ranks must be in 0..1024");
        ix->fast_front = int(value);
    } else if (n == "base_bytes") {
        ix->base_bytes = value != 0;
    } else if (n == "coarse_mfma") {
        ix->coarse_mfma = value != 0;
    } else if (n == "list_major") {
        ix->list_major = value != 0;
    } else if (n == "list_major_min_queries") {
        MIVF_REQUIRE(value >= 1, kInvalid, "list_major_min_queries must be >= 1");
        ix->list_major_min = value;
    } else if (n == "serve_min_fill") {
        MIVF_REQUIRE(value >= 0 && value <= (1 << 20), kInvalid, "serve_min_fill must be in 0..2^20");
        ix->serve_min_fill = value;
    } else if (n == "serve_max_wait_us") {
        MIVF_REQUIRE(value >= 0 && value <= 10000000, kInvalid, "serve_max_wait_us must be in 0..10^7");
        ix->serve_max_wait_us = value;
    } else if (n == "fast_capacity") {
        MIVF_REQUIRE(value >= 1024 && value <= (1 << 20), kInvalid, "fast_capacity must be in 1024..2^20");
        ix->fast_cap = int(value);
    } else {
        set_error("set_option: unknown option " + n);
        return kInvalid;
    }
    return kOk;
} MIVF_CATCH_ALL

int mivf_get_timings(mivf_index *ix, float *ms, int n) try {
    MIVF_TRY(check_index(ix));
    MIVF_REQUIRE(ms && n >= 0, kInvalid, "get_timings: bad arguments");
    for (int i = 0; i < n && i < 8; ++i) ms[i] = ix->timings[i];
    return kOk;
} MIVF_CATCH_ALL

int mivf_synth_rows(uint64_t seed, uint64_t stream_id, int64_t row0, int64_t n, int d, const float *d_centres,
                    int ncomp, const uint64_t *d_thr, float sigma, float *d_out, void *stream) try {
    MIVF_REQUIRE(n >= 0 && d > 0 && ncomp > 0 && d_centres && d_out, kInvalid, "synth_rows: bad arguments");
    MIVF_CHECK_HIP(launch_gen_rows(seed, stream_id, row0, n, d, d_centres, ncomp, d_thr, sigma, d_out,
                                   static_cast<hipStream_t>(stream)));
    return kOk;
} MIVF_CATCH_ALL

int mivf_synth_centres(uint64_t seed, int ncomp, int d, float *centres) try {
    MIVF_REQUIRE(ncomp > 0 && d > 0 && centres, kInvalid, "synth_centres: bad arguments");
    const uint64_t b = synth_stream_base(seed, 1);
    for (int64_t i = 0; i < int64_t(ncomp) * d; ++i) {
        uint64_t u = synth_draw(b, uint64_t(i)) >> 11;
        centres[i] = float(double(u) * (1.0 / 9007199254740992.0) * 128.0);
    }
    return kOk;
} MIVF_CATCH_ALL

int mivf_zipf_thresholds(uint64_t seed, int ncomp, double s, uint64_t *thr) try {
    MIVF_REQUIRE(ncomp > 0 && thr, kInvalid, "zipf_thresholds: bad arguments");
    std::vector<int> perm(ncomp), rank(ncomp);
    const uint64_t b = synth_stream_base(seed, 0x7a697066ull);
    for (int i = 0; i < ncomp; ++i) perm[i] = i;
    for (int i = 0; i < ncomp - 1; ++i) {
        uint64_t r = synth_draw(b, uint64_t(i));
        int j = i + int((unsigned __int128)r * uint64_t(ncomp - i) >> 64);
        std::swap(perm[i], perm[j]);
    }
    for (int i = 0; i < ncomp; ++i) rank[perm[i]] = i;
    double tot = 0;
    for (int c = 0; c < ncomp; ++c) tot += 1.0 / std::pow(double(rank[c] + 1), s);
    double cum = 0;
    for (int c = 0; c < ncomp; ++c) {
        cum += 1.0 / std::pow(double(rank[c] + 1), s);
        double t = std::floor(cum / tot * 9007199254740992.0);
        thr[c] = (c == ncomp - 1 || t > 9007199254740992.0) ? (uint64_t(1) << 53) : uint64_t(t);
    }
    return kOk;
} MIVF_CATCH_ALL

int mivf_bruteforce_nn(const float *d_base, int64_t nb, const float *d_q, int nq, int d, int32_t *d_ids,
                       double *d_dis) try {
    MIVF_REQUIRE(d == 128, kUnsupported, "bruteforce_nn: d = 128 only");
    MIVF_REQUIRE(nb > 0 && nq >= 0 && d_base && d_q && d_ids && d_dis, kInvalid, "bruteforce_nn: bad arguments");
    if (nq == 0) return kOk;
    const int nchunks = int(std::min<int64_t>(nb, 512));
    DevBuf pd, pi;
    MIVF_TRY(pd.ensure(sizeof(double) * size_t(nchunks) * nq));
    MIVF_TRY(pi.ensure(sizeof(int32_t) * size_t(nchunks) * nq));
    MIVF_CHECK_HIP(launch_bruteforce_nn(d_base, nb, d_q, nq, d, d_ids, d_dis, pd.as<double>(), pi.as<int32_t>(),
                                        nchunks, nullptr));
    MIVF_CHECK_HIP(hipDeviceSynchronize());
    return kOk;
} MIVF_CATCH_ALL

}  // extern "C"
