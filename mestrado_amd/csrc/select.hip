// Row top-k (coarse probe selection), the merge of per-group partial top-k
// lists, and the exact rerank of a shortlist.  All selections order by
// (distance ascending, id ascending) through 64-bit keys and are exact.
#include "select_dev.h"
#include "rerank_dev.h"

namespace mivf {

namespace {

constexpr int NT = 256;
constexpr int KBUF = 1024;  // max k

// Collect the selected keys into buf (unsorted), return count.
template <typename KeyFn>
__device__ int collect(KeyFn keyfn, int n, SelectResult r, uint64_t *buf, int *cnt, int *eqt) {
    const int tid = threadIdx.x;
    if (tid == 0) {
        *cnt = 0;
        *eqt = 0;
    }
    __syncthreads();
    for (int i = tid; i < n; i += NT) {
        uint64_t key = keyfn(i);
        if (key == kNoKey) continue;
        bool take = key < r.thr;
        if (!take && r.eq > 0 && key == r.thr) take = atomicAdd(eqt, 1) < r.eq;
        if (take) {
            const int pos = atomicAdd(cnt, 1);
            if (pos < KBUF) buf[pos] = key;
        }
    }
    __syncthreads();
    return min(*cnt, KBUF);
}

// buf[0..c) holds the selected keys (complete: the caller's barrier).
__device__ void sort_and_emit(uint64_t *buf, int c, int k, int32_t *idx, float *dis) {
    const int tid = threadIdx.x;
    if (c <= NT) {
        // key i goes to its rank: the count of smaller keys, plus equal keys
        // before it; every thread reads the keys in the same order
        if (tid < c) {
            const uint64_t key = buf[tid];
            int rank = 0;
            for (int j = 0; j < c; ++j) {
                const uint64_t o = buf[j];
                rank += o < key || (o == key && j < tid);
            }
            if (rank < k) {
                idx[rank] = key_id(key);
                dis[rank] = key_dist(key);
            }
        }
        for (int i = c + tid; i < k; i += NT) {
            idx[i] = -1;
            dis[i] = __builtin_huge_valf();
        }
        return;
    }
    const int np = next_pow2(c > 0 ? c : 1);
    for (int i = c + tid; i < np; i += NT) buf[i] = kNoKey;
    __syncthreads();
    block_bitonic_sort<NT>(buf, np);
    for (int i = tid; i < k; i += NT) {
        if (i < c) {
            idx[i] = key_id(buf[i]);
            dis[i] = key_dist(buf[i]);
        } else {
            idx[i] = -1;
            dis[i] = __builtin_huge_valf();
        }
    }
}

__global__ __launch_bounds__(NT) void row_topk_kernel(const float *__restrict__ D, int n,
                                                      int64_t ld, int k,
                                                      int32_t *__restrict__ idx,
                                                      float *__restrict__ dis) {
    __shared__ SelectShared sh;
    __shared__ uint64_t buf[KBUF];
    __shared__ int cnt, eqt;
    const int64_t row = blockIdx.x;
    const float *Dr = D + row * ld;
    auto keyfn = [&](int i) { return ckey(Dr[i], i); };
    const int kk = k < n ? k : n;
    SelectResult r = block_select<NT>(keyfn, n, kk, sh);
    int c = collect(keyfn, n, r, buf, &cnt, &eqt);
    sort_and_emit(buf, c, k, idx + row * k, dis + row * k);
}

// k smallest (distance, id) pairs of one candidate array, ascending: the
// selection of parallel/ivf_pq/k_min.cpp:3-73 done exactly.
__global__ __launch_bounds__(NT) void kmin_kernel(const float *__restrict__ dis, const int32_t *__restrict__ ids,
                                                  int n, int k, int32_t *__restrict__ out_ids,
                                                  float *__restrict__ out_dis) {
    __shared__ SelectShared sh;
    __shared__ uint64_t buf[KBUF];
    __shared__ int cnt, eqt;
    auto keyfn = [&](int i) { return ckey(dis[i], ids[i]); };
    const int kk = k < n ? k : n;
    SelectResult r = block_select<NT>(keyfn, n, kk, sh);
    int c = collect(keyfn, n, r, buf, &cnt, &eqt);
    sort_and_emit(buf, c, k, out_ids, out_dis);
}

__global__ __launch_bounds__(NT) void merge_kernel(const uint64_t *__restrict__ partial,
                                                   int ngroups, int kstride,
                                                   const int32_t *__restrict__ kq, int k,
                                                   int32_t *__restrict__ out_ids,
                                                   float *__restrict__ out_dis) {
    __shared__ SelectShared sh;
    __shared__ uint64_t buf[KBUF];
    __shared__ int cnt, eqt;
    const int q = blockIdx.x;
    const uint64_t *P = partial + int64_t(q) * ngroups * kstride;
    const int n = ngroups * kstride;
    const int kk = kq ? kq[q] : k;
    auto keyfn = [&](int i) { return P[i]; };
    SelectResult r = block_select<NT>(keyfn, n, kk, sh);
    int c = collect(keyfn, n, r, buf, &cnt, &eqt);
    // rows are kstride wide; entries past this query's k are padding
    sort_and_emit(buf, c, kstride, out_ids + int64_t(q) * kstride, out_dis + int64_t(q) * kstride);
}

// rerank: a = base row (f32-accumulated norm), b = query (f64-accumulated norm),
// as knn_full(2, 1, ktmp, d, k, nearest, query) in sequential/ivf_pq/ivf_search.cpp:130.
// MODE 0: the shortlist holds ids, rows base[id].  MODE 1: it holds list
// positions, id = list_ids[pos], rows base[id].  MODE 2: list positions, rows
// from the list-ordered copy base[pos] (nbase = its row count).  qnorm: the
// queries' f64-accumulated norms when the caller has them (the same sum as
// below), else nullptr.  T: float rows, or byte rows of integer-valued data
// (the same float values, so the same distances bit for bit).
template <int MODE, typename T = float>
__global__ __launch_bounds__(NT) void rerank_kernel(const T *__restrict__ base, int64_t nbase,
                                                    const int32_t *__restrict__ list_ids,
                                                    const float *__restrict__ queries, int d,
                                                    const double *__restrict__ qnorm,
                                                    const int32_t *__restrict__ short_ids,
                                                    int kl, int k, int32_t *__restrict__ out_ids,
                                                    float *__restrict__ out_dis) {
    __shared__ SelectShared sh;
    __shared__ uint64_t cand[KBUF];
    __shared__ uint64_t buf[KBUF];
    __shared__ float qv[512];
    __shared__ double qn;
    __shared__ int cnt, eqt;
    const int tid = threadIdx.x;
    const int q = blockIdx.x;
    for (int t = tid; t < d; t += NT) qv[t] = queries[int64_t(q) * d + t];
    __syncthreads();
    if (tid == 0) {
        if (qnorm) {
            qn = qnorm[q];
        } else {
            double s = 0.0;
            for (int t = 0; t < d; ++t) {
                float p = qv[t] * qv[t];
                s = s + double(p);
            }
            qn = s;
        }
    }
    __syncthreads();
    const double nb = qn;
    const bool vec = (d % (64 / int(sizeof(T)))) == 0 && (reinterpret_cast<uintptr_t>(base) & 15u) == 0;
    for (int i = tid; i < kl; i += NT) {
        const int32_t e = short_ids[int64_t(q) * kl + i];
        int32_t id = e;
        int64_t row = e;
        if (MODE != 0 && e >= 0) id = list_ids[e];
        if (MODE == 1) row = id;
        uint64_t key = kNoKey;
        if (e >= 0 && id >= 0 && row < nbase) {
            float na = 0.0f, dot = 0.0f;
            row_sums(base + row * d, qv, d, vec, na, dot);
            float c = float(nb + double(na));
            float m2 = -2.0f * dot;
            key = ckey(m2 + c, id);
        }
        cand[i] = key;
    }
    __syncthreads();
    auto keyfn = [&](int i) { return cand[i]; };
    SelectResult r = block_select<NT>(keyfn, kl, k, sh);
    int c = collect(keyfn, kl, r, buf, &cnt, &eqt);
    sort_and_emit(buf, c, k, out_ids + int64_t(q) * k, out_dis + int64_t(q) * k);
}

// The rerank for shortlists of at most NT (one row per thread): rerank_block
// (rerank_dev.h) with the shortlist entry read from short_ids.
template <int MODE, typename T = float>
__global__ __launch_bounds__(NT) void rerank_small_kernel(const T *__restrict__ base, int64_t nbase,
                                                          const int32_t *__restrict__ list_ids,
                                                          const float *__restrict__ queries, int d,
                                                          const double *__restrict__ qnorm,
                                                          const int32_t *__restrict__ short_ids,
                                                          int kl, int k, int32_t *__restrict__ out_ids,
                                                          float *__restrict__ out_dis) {
    __shared__ RerankShared rs;
    const int q = blockIdx.x;
    const int32_t e = int(threadIdx.x) < kl ? short_ids[int64_t(q) * kl + threadIdx.x] : -1;
    rerank_block<MODE, T, NT>(base, nbase, list_ids, queries, d, qnorm, q, e, k, out_ids, out_dis, rs);
}

__global__ void gather_rows_kernel(const float *__restrict__ base, int d, const int32_t *__restrict__ ids,
                                   int64_t n, float *__restrict__ out) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * d) return;
    const int64_t r = e / d;
    out[e] = base[int64_t(ids[r]) * d + (e - r * d)];
}

// Byte copy of the rows in list order; *bad is set when a value is not an
// integer in 0..255 (the copy would not be exact).
__global__ void gather_rows_u8_kernel(const float *__restrict__ base, int d, const int32_t *__restrict__ ids,
                                      int64_t n, uint8_t *__restrict__ out, int *__restrict__ bad) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * d) return;
    const int64_t r = e / d;
    const float v = base[int64_t(ids[r]) * d + (e - r * d)];
    const bool ok = v >= 0.0f && v <= 255.0f && v == __builtin_truncf(v);
    if (!ok) atomicOr(bad, 1);
    out[e] = ok ? uint8_t(v) : uint8_t(0);
}

}  // namespace

hipError_t launch_row_topk(const float *D, int64_t rows, int n, int64_t ld, int k,
                           int32_t *idx, float *dis, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    if (k < 1 || k > KBUF) return hipErrorInvalidValue;
    row_topk_kernel<<<unsigned(rows), NT, 0, s>>>(D, n, ld, k, idx, dis);
    return hipGetLastError();
}

hipError_t launch_kmin(const float *dis, const int32_t *ids, int64_t n, int k, int32_t *out_ids, float *out_dis,
                       hipStream_t s) {
    if (k < 1 || k > KBUF || n < 0 || n > INT32_MAX) return hipErrorInvalidValue;
    kmin_kernel<<<1, NT, 0, s>>>(dis, ids, int(n), k, out_ids, out_dis);
    return hipGetLastError();
}

hipError_t launch_merge(const uint64_t *partial, int nq, int ngroups, int kstride,
                        const int32_t *kq, int k, int32_t *out_ids, float *out_dis,
                        hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    if (kstride < 1 || kstride > KBUF) return hipErrorInvalidValue;
    merge_kernel<<<unsigned(nq), NT, 0, s>>>(partial, ngroups, kstride, kq, k, out_ids, out_dis);
    return hipGetLastError();
}

hipError_t launch_rerank(const float *base, int64_t nbase, const float *queries, int nq,
                         int d, const int32_t *short_ids, int kl, int k,
                         int32_t *out_ids, float *out_dis, hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    if (kl < 1 || kl > KBUF || k < 1 || k > KBUF || d > 512) return hipErrorInvalidValue;
    if (kl <= NT)
        rerank_small_kernel<0><<<unsigned(nq), NT, 0, s>>>(base, nbase, nullptr, queries, d, nullptr, short_ids, kl,
                                                           k, out_ids, out_dis);
    else
        rerank_kernel<0><<<unsigned(nq), NT, 0, s>>>(base, nbase, nullptr, queries, d, nullptr, short_ids, kl, k,
                                                     out_ids, out_dis);
    return hipGetLastError();
}

hipError_t launch_rerank_pos(const float *rows, int64_t nrows, bool list_ordered, const int32_t *list_ids,
                             const float *queries, int nq, int d, const double *qnorm, const int32_t *short_pos,
                             int kl, int k, int32_t *out_ids, float *out_dis, hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    if (kl < 1 || kl > KBUF || k < 1 || k > KBUF || d > 512) return hipErrorInvalidValue;
    if (list_ordered && kl <= NT)
        rerank_small_kernel<2><<<unsigned(nq), NT, 0, s>>>(rows, nrows, list_ids, queries, d, qnorm, short_pos, kl,
                                                           k, out_ids, out_dis);
    else if (list_ordered)
        rerank_kernel<2><<<unsigned(nq), NT, 0, s>>>(rows, nrows, list_ids, queries, d, qnorm, short_pos, kl, k,
                                                     out_ids, out_dis);
    else if (kl <= NT)
        rerank_small_kernel<1><<<unsigned(nq), NT, 0, s>>>(rows, nrows, list_ids, queries, d, qnorm, short_pos, kl,
                                                           k, out_ids, out_dis);
    else
        rerank_kernel<1><<<unsigned(nq), NT, 0, s>>>(rows, nrows, list_ids, queries, d, qnorm, short_pos, kl, k,
                                                     out_ids, out_dis);
    return hipGetLastError();
}

hipError_t launch_rerank_pos_u8(const uint8_t *rows, int64_t nrows, const int32_t *list_ids, const float *queries,
                                int nq, int d, const double *qnorm, const int32_t *short_pos, int kl, int k,
                                int32_t *out_ids, float *out_dis, hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    if (kl < 1 || kl > KBUF || k < 1 || k > KBUF || d > 512) return hipErrorInvalidValue;
    if (kl <= NT)
        rerank_small_kernel<2, uint8_t><<<unsigned(nq), NT, 0, s>>>(rows, nrows, list_ids, queries, d, qnorm,
                                                                    short_pos, kl, k, out_ids, out_dis);
    else
        rerank_kernel<2, uint8_t><<<unsigned(nq), NT, 0, s>>>(rows, nrows, list_ids, queries, d, qnorm, short_pos,
                                                              kl, k, out_ids, out_dis);
    return hipGetLastError();
}

hipError_t launch_gather_rows_u8(const float *base, int d, const int32_t *ids, int64_t n, uint8_t *out, int *bad,
                                 hipStream_t s) {
    if (n <= 0) return hipSuccess;
    gather_rows_u8_kernel<<<unsigned((n * d + 255) / 256), 256, 0, s>>>(base, d, ids, n, out, bad);
    return hipGetLastError();
}

hipError_t launch_gather_rows(const float *base, int d, const int32_t *ids, int64_t n, float *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    gather_rows_kernel<<<unsigned((n * d + 255) / 256), 256, 0, s>>>(base, d, ids, n, out);
    return hipGetLastError();
}

}  // namespace mivf
