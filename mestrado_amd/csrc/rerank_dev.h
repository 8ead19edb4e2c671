// Exact FP32 rerank of shortlisted raw rows: the distance arithmetic of
// knn_full(2, 1, ktmp, d, k, nearest, query) in
// sequential/ivf_pq/ivf_search.cpp:130 (a = base row with its f32-accumulated
// norm, b = query with its f64-accumulated norm), shared by the rerank
// kernels of select.hip and the fused shortlist + rerank of fast.hip.
#pragma once
#include "select_dev.h"

namespace mivf {
namespace {

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

struct RerankShared {
    uint64_t cand[256];
    float qv[512];
    uint32_t qb[32];  // the query as bytes (d = 128, integer-valued queries)
    double qn;
};

// The exact rerank of a shortlist of at most NT entries, one per thread
// (e: this thread's entry, -1 for none), by the whole workgroup.  The query
// and its norm are loaded while the entry arrives, a 128-byte
// row (bytes, d = 128) is loaded as soon as its position is known (MODE 0
// and 2: the row is the entry itself) while the id is still on its way, and
// each key then goes straight to its rank (the count of smaller keys, equal
// keys ordered by position), found from per-wavefront sorted runs.
// Byte rows against a query whose values are all integers in 0..255 (checked
// per query): every product and partial sum of |x|^2 and x.q is then an
// integer below 128 * 255^2 < 2^24, so the sequential FP32 sums are exact
// and equal the integer sums, which v_dot4_u32_u8 forms four bytes at a time.
template <int MODE, typename T, int NT>
__device__ void rerank_block(const T *__restrict__ base, int64_t nbase, const int32_t *__restrict__ list_ids,
                             const float *__restrict__ queries, int d, const double *__restrict__ qnorm, int q,
                             int32_t e, int k, int32_t *__restrict__ out_ids, float *__restrict__ out_dis,
                             RerankShared &rs) {
    static_assert(NT <= 256 && NT % 64 == 0, "one shortlist entry per thread, at most 256, whole wavefronts");
    uint64_t *cand = rs.cand;
    float *qv = rs.qv;
    uint32_t *qb = rs.qb;
    double &qn = rs.qn;
    const int tid = threadIdx.x;
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
    // Rank of each key among the workgroup's keys, equal keys ordered by
    // thread: every wavefront sorts its 64 (key, thread) pairs (bitonic
    // network over lanes), publishes the sorted run, and the pair at lane L
    // of run w then has rank L + the count of smaller pairs in the other
    // runs (binary searches: a key of run w2 < w comes first when equal).
    const int lane = tid & 63, wave = tid >> 6;
    uint64_t sk = key;
    int sj = tid;
#pragma unroll
    for (int kk = 2; kk <= 64; kk <<= 1) {
#pragma unroll
        for (int j = kk >> 1; j > 0; j >>= 1) {
            const uint32_t plo = uint32_t(__shfl_xor(int(uint32_t(sk)), j, 64));
            const uint32_t phi = uint32_t(__shfl_xor(int(uint32_t(sk >> 32)), j, 64));
            const uint64_t pk = (uint64_t(phi) << 32) | plo;
            const int pj = __shfl_xor(sj, j, 64);
            const bool pless = pk < sk || (pk == sk && pj < sj);
            // ascending blocks keep the smaller pair in the lower lane, descending ones the larger
            const bool take = (((lane & kk) == 0) == ((lane & j) == 0)) ? pless : !pless;
            if (take) {
                sk = pk;
                sj = pj;
            }
        }
    }
    cand[tid] = sk;  // run `wave`, ascending
    const int nreal = __syncthreads_count(key != kNoKey);
    int32_t *oi = out_ids + int64_t(q) * k;
    float *od = out_dis + int64_t(q) * k;
    if (sk != kNoKey) {
        int rank = lane;
#pragma unroll
        for (int w2 = 0; w2 < NT / 64; ++w2) {
            if (w2 == wave) continue;
            const uint64_t *run = cand + 64 * w2;
            const bool le = w2 < wave;  // equal keys of an earlier run come first
            int pos = 0;
#pragma unroll
            for (int s = 32; s > 0; s >>= 1) {
                const uint64_t o = run[pos + s - 1];
                if (le ? o <= sk : o < sk) pos += s;
            }
            const uint64_t o = run[pos];
            pos += (le ? o <= sk : o < sk) ? 1 : 0;
            rank += pos;
        }
        if (rank < k) {
            oi[rank] = key_id(sk);
            od[rank] = key_dist(sk);
        }
    }
    for (int i = nreal + tid; i < k; i += NT) {
        oi[i] = -1;
        od[i] = __builtin_huge_valf();
    }
}

}  // namespace
}  // namespace mivf
