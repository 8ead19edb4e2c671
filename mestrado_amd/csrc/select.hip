// Row top-k (coarse probe selection), the merge of per-group partial top-k
// lists, and the exact rerank of a shortlist.  All selections order by
// (distance ascending, id ascending) through 64-bit keys and are exact.
#include "select_dev.h"

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

// Element t of row x as a float: rows of float, or of bytes holding the
// integers 0..255 exactly (a byte row converts to the same floats).
__device__ inline float row_el(const float *x, int t) { return x[t]; }
__device__ inline float row_el(const uint8_t *x, int t) { return float(x[t]); }

// The squared norm (f32 accumulator) and the dot product with qv of one row,
// both sequential in element order with every product rounded (no FMA).
__device__ inline void row_sums(const float *x, const float *qv, int d, bool vec, float &na, float &dot) {
    int t = 0;
    if (vec) {
        // 16-byte loads, issued 8 at a time; the sums keep the element order
        const float4 *x4 = reinterpret_cast<const float4 *>(x);
        for (; t + 32 <= d; t += 32) {
            float4 v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = x4[t / 4 + j];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float el[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    na = na + el[u] * el[u];
                    dot = dot + el[u] * qv[t + 4 * j + u];
                }
            }
        }
    }
    for (; t < d; ++t) {
        const float xv = x[t];
        const float p = xv * xv;
        na = na + p;
        const float pd = xv * qv[t];
        dot = dot + pd;
    }
}

__device__ inline void row_sums(const uint8_t *x, const float *qv, int d, bool vec, float &na, float &dot) {
    int t = 0;
    if (vec) {
        // one 128-byte row in eight 16-byte loads issued together
        const uint4 *x16 = reinterpret_cast<const uint4 *>(x);
        for (; t + 128 <= d; t += 128) {
            uint4 v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = x16[t / 16 + j];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t wd[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const float el = float((wd[u >> 2] >> (8 * (u & 3))) & 255u);
                    na = na + el * el;
                    dot = dot + el * qv[t + 16 * j + u];
                }
            }
        }
    }
    for (; t < d; ++t) {
        const float xv = row_el(x, t);
        const float p = xv * xv;
        na = na + p;
        const float pd = xv * qv[t];
        dot = dot + pd;
    }
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

// The same rerank for shortlists of at most NT (one row per thread): the
// shortlist entry, the query and its norm are loaded together, a 128-byte
// row (bytes, d = 128) is loaded as soon as its position is known (MODE 0
// and 2: the row is the entry itself) while the id is still on its way, and
// each key then goes straight to its rank (the count of smaller keys, equal
// keys ordered by position).
// Byte rows against a query whose values are all integers in 0..255 (checked
// per query): every product and partial sum of |x|^2 and x.q is then an
// integer below 128 * 255^2 < 2^24, so the sequential FP32 sums are exact
// and equal the integer sums, which v_dot4_u32_u8 forms four bytes at a time.
template <int MODE, typename T = float>
__global__ __launch_bounds__(NT) void rerank_small_kernel(const T *__restrict__ base, int64_t nbase,
                                                          const int32_t *__restrict__ list_ids,
                                                          const float *__restrict__ queries, int d,
                                                          const double *__restrict__ qnorm,
                                                          const int32_t *__restrict__ short_ids,
                                                          int kl, int k, int32_t *__restrict__ out_ids,
                                                          float *__restrict__ out_dis) {
    __shared__ uint64_t cand[NT];
    __shared__ float qv[512];
    __shared__ uint32_t qb[32];  // the query as bytes (d = 128, integer-valued queries)
    __shared__ double qn;
    const int tid = threadIdx.x;
    const int q = blockIdx.x;
    const int32_t e = tid < kl ? short_ids[int64_t(q) * kl + tid] : -1;
    bool qint = true;
    for (int t = tid; t < d; t += NT) {
        const float v = queries[int64_t(q) * d + t];
        qv[t] = v;
        qint = qint && v >= 0.0f && v <= 255.0f && v == __builtin_truncf(v);
        if (t < 128) reinterpret_cast<uint8_t *>(qb)[t] = uint8_t(qint ? v : 0.0f);
    }
    if (tid == 0 && qnorm) qn = qnorm[q];
    // whole byte rows of 128 values in registers (eight 16-byte loads)
    constexpr bool kRow = sizeof(T) == 1 && MODE != 1;
    const bool whole = kRow && d == 128 && (reinterpret_cast<uintptr_t>(base) & 15u) == 0;
    uint4 pre[8];
    if (whole && e >= 0 && int64_t(e) < nbase) {
        const uint4 *x16 = reinterpret_cast<const uint4 *>(reinterpret_cast<const uint8_t *>(base) + int64_t(e) * 128);
#pragma unroll
        for (int j = 0; j < 8; ++j) pre[j] = x16[j];
    }
    int32_t id = e;
    if (MODE != 0 && e >= 0) id = list_ids[e];
    const int64_t row = MODE == 1 ? int64_t(id) : int64_t(e);
    const bool ibytes = __syncthreads_and(qint) && whole;  // qv, qb and qn are ready
    if (!qnorm) {
        if (tid == 0) {
            double s = 0.0;
            for (int t = 0; t < d; ++t) {
                const float p = qv[t] * qv[t];
                s = s + double(p);
            }
            qn = s;
        }
        __syncthreads();
    }
    uint64_t key = kNoKey;
    if (e >= 0 && id >= 0 && row < nbase) {
        float na = 0.0f, dot = 0.0f;
        if (ibytes) {
            uint32_t ina = 0u, idot = 0u;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t wd[4] = {pre[j].x, pre[j].y, pre[j].z, pre[j].w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    ina = __builtin_amdgcn_udot4(wd[u], wd[u], ina, false);
                    idot = __builtin_amdgcn_udot4(wd[u], qb[4 * j + u], idot, false);
                }
            }
            na = float(ina);
            dot = float(idot);
        } else if (whole) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t wd[4] = {pre[j].x, pre[j].y, pre[j].z, pre[j].w};
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const float el = float((wd[u >> 2] >> (8 * (u & 3))) & 255u);
                    na = na + el * el;
                    dot = dot + el * qv[16 * j + u];
                }
            }
        } else {
            const bool vec = (d % (64 / int(sizeof(T)))) == 0 && (reinterpret_cast<uintptr_t>(base) & 15u) == 0;
            row_sums(base + row * d, qv, d, vec, na, dot);
        }
        const float c = float(qn + double(na));
        const float m2 = -2.0f * dot;
        key = ckey(m2 + c, id);
    }
    cand[tid] = key;
    const int nreal = __syncthreads_count(key != kNoKey);
    int32_t *oi = out_ids + int64_t(q) * k;
    float *od = out_dis + int64_t(q) * k;
    if (key != kNoKey) {
        int rank = 0;
        for (int j = 0; j < kl; ++j) {
            const uint64_t o = cand[j];
            rank += o < key || (o == key && j < tid);
        }
        if (rank < k) {
            oi[rank] = key_id(key);
            od[rank] = key_dist(key);
        }
    }
    for (int i = nreal + tid; i < k; i += NT) {
        oi[i] = -1;
        od[i] = __builtin_huge_valf();
    }
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
