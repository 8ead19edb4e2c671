// Batched reduced-precision IVF-PQ search (approximate scan + exact rerank).
//
// Work is list-major: a work item is (list l, up to FG = 16 queries that probe
// l).  Its 16 distance tables are quantized to 8 bits and interleaved in LDS
// as t8[s][code][query], so one ds_read_b128 gathers the entries of all 16
// queries for a vector's code (the code belongs to the vector, not to the
// query), and the list's codes are read from HBM once per 16 queries.
//
// Table entry (query g, subquantizer s, code c), with r = q - c_l:
//   T = |r_s|^2 + (|p_sc|^2 + 2 c_ls.p_sc) - 2 q_s.p_sc          (= |r_s - p_sc|^2)
//   t8 = round(T * 255 / bound_g), bound_g = max_s (|r_s| + max_c |p_sc|)^2 >= T
// Approximate distance = delta_g * sum_s t8 with delta_g = bound_g / 255; each
// term is within delta_g / 2 of T (plus FP32 rounding), so
// |approx - ADC| <= 8 delta_g (+ slack) for m = 16.
//
// Candidate filter: a per-query threshold tau_q from a seed pass (the exact
// ADC kl-th distance over the query's nearest list).  A vector passes if
// sum <= tau_q / delta_g + m/2 + 2, which keeps every vector whose exact ADC
// distance is <= tau_q.  Passing (approx key, id) pairs are appended to a
// per-query global buffer; fast_select keeps the kl best (approx distance, id)
// together with the seed's, and the exact FP32 rerank orders them.
#include "select_dev.h"

namespace mivf {

namespace {

constexpr int D = 128;
constexpr int FG = 16;    // queries per work item
constexpr int FNT = 256;  // threads per workgroup

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const u32x4 lds_u32x4;

// Candidate staging in LDS: entries (list position v, query slot g, sum) are
// appended with one LDS atomic per wavefront and flushed to the per-query
// global buffers with one global atomic per query slot.
constexpr int POOL = 1536;

template <int M>
struct FLay {
    static constexpr size_t t8 = 0;                            // uint8 [M][256][FG]
    static constexpr size_t pool = t8 + size_t(M) * 256 * FG;  // uint2 [POOL] (scan phase)
    static constexpr size_t qv = pool;                         // float [FG][D]   (table phase)
    static constexpr size_t A = qv + size_t(FG) * D * 4;       // float [FG][M]   (table phase)
    static constexpr size_t cl = A + size_t(FG) * M * 4;       // float [D]       (table phase)
    static_assert(cl + D * 4 <= pool + POOL * 8, "table-phase data must fit in the pool");
    static constexpr size_t invd = pool + POOL * 8;            // float [FG]
    static constexpr size_t delt = invd + FG * 4;              // float [FG]
    static constexpr size_t sthr = delt + FG * 4;              // uint32 [FG/2] (u16 pairs of limits)
    static constexpr size_t stmp = sthr + FG / 2 * 4;          // uint32 [FG] unpacked limits
    static constexpr size_t qid = stmp + FG * 4;               // int [FG]
    static constexpr size_t gcnt = qid + FG * 4;               // int [FG]
    static constexpr size_t gbase = gcnt + FG * 4;             // int [FG]
    static constexpr size_t pcnt = gbase + FG * 4;             // int
    static constexpr size_t item = pcnt + 4;                   // int
    static constexpr size_t total = item + 12;
};

// Move the staged candidates to the per-query global buffers (whole
// workgroup, uniform call).  Entry: x = list position, y = g | sum << 4,
// and the rank within its query slot goes into bits 16..27 of y.
template <int M>
__device__ void flush_pool(const FastArgs &a, unsigned char *smem, int64_t beg) {
    using L = FLay<M>;
    uint2 *pool = reinterpret_cast<uint2 *>(smem + L::pool);
    int *gcnt = reinterpret_cast<int *>(smem + L::gcnt);
    int *gbase = reinterpret_cast<int *>(smem + L::gbase);
    int *pcnt = reinterpret_cast<int *>(smem + L::pcnt);
    const int *qid = reinterpret_cast<const int *>(smem + L::qid);
    const float *delt = reinterpret_cast<const float *>(smem + L::delt);
    const int tid = threadIdx.x;
    const int n = min(*pcnt, POOL);
    if (tid < FG) gcnt[tid] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += FNT) {
        const uint32_t y = pool[i].y;
        const int r = atomicAdd(gcnt + (y & 15u), 1);
        pool[i].y = y | (uint32_t(r) << 16);
    }
    __syncthreads();
    if (tid < FG && gcnt[tid] > 0) gbase[tid] = atomicAdd(a.cand_cnt + qid[tid], gcnt[tid]);
    __syncthreads();
    for (int i = tid; i < n; i += FNT) {
        const uint2 e = pool[i];
        const int g = int(e.y & 15u);
        const uint32_t sum = (e.y >> 4) & 0xFFFu;
        const int pos = gbase[g] + int(e.y >> 16);
        if (pos < a.capq)
            a.cand[int64_t(qid[g]) * a.capq + pos] = ckey(delt[g] * float(sum), a.list_ids[beg + e.x]);
    }
    __syncthreads();
    if (tid == 0) *pcnt = 0;
    __syncthreads();
}

template <int M>
__device__ inline void load_codes(const uint8_t *lc, int v, uint32_t (&cw)[M / 4]) {
    if constexpr (M == 16) {
        const uint4 c4 = *reinterpret_cast<const uint4 *>(lc + uint32_t(v) * 16u);
        cw[0] = c4.x; cw[1] = c4.y; cw[2] = c4.z; cw[3] = c4.w;
    } else {
        const uint2 c2 = *reinterpret_cast<const uint2 *>(lc + uint32_t(v) * 8u);
        cw[0] = c2.x; cw[1] = c2.y;
    }
}

// Accumulators: ae[i] holds the sums of queries 4i (low u16) and 4i+2 (high),
// ao[i] those of 4i+1 and 4i+3; per query the sum is <= m * 255 < 2^16, so
// the low half never carries into the high one.
template <int M>
__global__ __launch_bounds__(FNT) void fast_scan_kernel(FastArgs a) {
    constexpr int DS = D / M;
    using L = FLay<M>;
    extern __shared__ __align__(16) unsigned char smem[];
    float *qv = reinterpret_cast<float *>(smem + L::qv);
    float *A = reinterpret_cast<float *>(smem + L::A);
    float *cl = reinterpret_cast<float *>(smem + L::cl);
    float *invd = reinterpret_cast<float *>(smem + L::invd);
    float *delt = reinterpret_cast<float *>(smem + L::delt);
    uint32_t *sthr = reinterpret_cast<uint32_t *>(smem + L::sthr);
    uint32_t *stmp = reinterpret_cast<uint32_t *>(smem + L::stmp);
    int *qid = reinterpret_cast<int *>(smem + L::qid);
    int *pcnt = reinterpret_cast<int *>(smem + L::pcnt);
    int *item_sh = reinterpret_cast<int *>(smem + L::item);
    uint2 *pool = reinterpret_cast<uint2 *>(smem + L::pool);
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    if (uint32_t(uintptr_t((__attribute__((address_space(3))) unsigned char *)(smem))) != 0) {
        if (tid == 0) atomicOr(a.err, 2);
        return;
    }
    const int total = *a.total_items;
    if (tid == 0) *pcnt = 0;

    for (;;) {
        if (tid == 0) *item_sh = atomicAdd(a.next_item, 1);
        __syncthreads();
        const int item = *item_sh;
        if (item >= total) break;  // uniform
        const int l = a.item_list[item];
        const int pb = a.item_pair[item];
        const int gc = min(FG, a.list_pair_off[l + 1] - pb);
        const int64_t beg = a.list_off[l];
        const int len = int(a.list_off[l + 1] - beg);

        if (tid < FG) qid[tid] = tid < gc ? a.pair_q[pb + tid] : -1;
        if (tid < D) cl[tid] = a.coarse[int64_t(l) * D + tid];
        __syncthreads();
        for (int e = tid; e < FG * D; e += FNT) {
            const int g = e / D;
            const int q = qid[g];
            qv[e] = q >= 0 ? a.queries[int64_t(q) * D + (e % D)] : 0.0f;
        }
        __syncthreads();
        {   // |r_s|^2 for (g, s): FG*M <= 256 threads
            const int g = tid / M, s = tid % M;
            if (g < FG) {
                float acc = 0.0f;
#pragma unroll
                for (int t = 0; t < DS; ++t) {
                    const float r = qv[g * D + s * DS + t] - cl[s * DS + t];
                    acc = __builtin_fmaf(r, r, acc);
                }
                A[g * M + s] = acc;
            }
        }
        __syncthreads();
        if (tid < FG) {
            const int g = tid;
            float bound = 0.0f;
            for (int s = 0; s < M; ++s) {
                const float b = __builtin_sqrtf(A[g * M + s]) + a.pmax[s];
                bound = fmaxf(bound, b * b);
            }
            bound = fmaxf(bound * 1.0001f, 1e-20f);
            const float dlt = bound / 255.0f;
            delt[g] = dlt;
            invd[g] = 255.0f / bound;
            const float tau = qid[g] >= 0 ? a.tau[qid[g]] : -1.0f;
            // stored limit = thr + 1 (a sum passes iff sum < limit)
            uint32_t lim;
            if (qid[g] < 0) lim = 0u;                     // empty slot: nothing passes
            else if (!(tau < 3.0e38f)) lim = 0xFFFFu;     // no seed bound: keep everything
            else lim = uint32_t(fminf(tau / dlt + float(M / 2 + 2), 65534.0f)) + 1u;
            stmp[g] = lim;
        }
        __syncthreads();
        if (tid < FG / 2) {
            // sthr[k] holds the limits of the two queries of accumulator k:
            // k = 2i + odd pairs queries 4i + odd (low half) and 4i + odd + 2 (high half)
            const int i = tid >> 1, odd = tid & 1;
            const int qa = 4 * i + odd, qb = qa + 2;
            sthr[tid] = stmp[qa] | (stmp[qb] << 16);
        }

        // ---- 16 quantized tables: thread c owns code c of every subquantizer
        {
            const int c = tid;  // FNT == 256 == ks
            const bool cok = c < a.ks;
            const float *src = a.pq + int64_t(cok ? c : 0) * DS;
            const int64_t sstride = int64_t(a.ks) * DS;
            float pnext[DS];
#pragma unroll
            for (int t = 0; t < DS; ++t) pnext[t] = src[t];
#pragma unroll 1
            for (int s = 0; s < M; ++s) {
                float p[DS];
#pragma unroll
                for (int t = 0; t < DS; ++t) p[t] = pnext[t];
                if (s + 1 < M) {
#pragma unroll
                    for (int t = 0; t < DS; ++t) pnext[t] = src[(s + 1) * sstride + t];
                }
                float pn = 0.0f, cp = 0.0f;
#pragma unroll
                for (int t = 0; t < DS; ++t) {
                    pn = __builtin_fmaf(p[t], p[t], pn);
                    cp = __builtin_fmaf(cl[s * DS + t], p[t], cp);
                }
                const float B = pn + 2.0f * cp;
                uint32_t w4[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int g = 0; g < FG; ++g) {
                    float qp = 0.0f;
#pragma unroll
                    for (int t = 0; t < DS; ++t) qp = __builtin_fmaf(qv[g * D + s * DS + t], p[t], qp);
                    const float T = fmaxf(A[g * M + s] + B - 2.0f * qp, 0.0f);
                    const float v = cok ? T * invd[g] : 255.0f;
                    w4[g >> 2] = __builtin_amdgcn_cvt_pk_u8_f32(v, uint32_t(g & 3), w4[g >> 2]);
                }
                uint4 o;
                o.x = w4[0];
                o.y = w4[1];
                o.z = w4[2];
                o.w = w4[3];
                *reinterpret_cast<uint4 *>(smem + L::t8 + (size_t(s) * 256 + c) * FG) = o;
            }
        }
        __syncthreads();  // tables ready; the pool may now overwrite qv, A, cl

        // ---- scan: each lane one vector per step, codes loaded one step ahead
        const uint8_t *lc = a.codes + beg * M;
        uint32_t th[FG / 2];
#pragma unroll
        for (int k = 0; k < FG / 2; ++k) th[k] = sthr[k];
        uint32_t cn[M / 4];
        load_codes<M>(lc, min(tid, max(len - 1, 0)), cn);
        for (int base = 0; base < len; base += FNT) {
            const int v = base + tid;
            uint32_t cw[M / 4];
#pragma unroll
            for (int i = 0; i < M / 4; ++i) cw[i] = cn[i];
            if (base + FNT < len) load_codes<M>(lc, min(v + FNT, len - 1), cn);
            uint32_t ae[4] = {0u, 0u, 0u, 0u}, ao[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int s = 0; s < M; ++s) {
                const uint32_t code = (cw[s >> 2] >> (8 * (s & 3))) & 255u;
                const u32x4 x = *reinterpret_cast<lds_u32x4 *>(uintptr_t(L::t8 + (uint32_t(s) * 256u + code) * FG));
                const uint32_t xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    ae[i] += xs[i] & 0x00FF00FFu;
                    ao[i] += __builtin_amdgcn_perm(0u, xs[i], 0x0C030C01u);
                }
            }
            // pass mask over the 16 query slots: slot g = 4i + 2hi + odd
            uint32_t pm = 0u;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t te = th[2 * i], to = th[2 * i + 1];
                pm |= uint32_t((ae[i] & 0xFFFFu) < (te & 0xFFFFu)) << (4 * i);
                pm |= uint32_t((ao[i] & 0xFFFFu) < (to & 0xFFFFu)) << (4 * i + 1);
                pm |= uint32_t((ae[i] >> 16) < (te >> 16)) << (4 * i + 2);
                pm |= uint32_t((ao[i] >> 16) < (to >> 16)) << (4 * i + 3);
            }
            if (v >= len) pm = 0u;
            if (__ballot(pm != 0u) != 0) {  // wave-uniform
                // wave prefix sum of the per-lane pass counts (<= 16: five bit planes)
                const uint32_t n = uint32_t(__popc(pm));
                uint32_t pre = 0u, tot = 0u;
#pragma unroll
                for (int b = 0; b < 5; ++b) {
                    const uint64_t bm = __ballot((n >> b) & 1u);
                    pre += uint32_t(__builtin_amdgcn_mbcnt_hi(uint32_t(bm >> 32),
                                                              __builtin_amdgcn_mbcnt_lo(uint32_t(bm), 0u))) << b;
                    tot += uint32_t(__popcll(bm)) << b;
                }
                int slot = 0;
                if (lane == 0) slot = atomicAdd(pcnt, int(tot));
                slot = __shfl(slot, 0) + int(pre);
                while (pm) {
                    const int g = __ffs(pm) - 1;
                    pm &= pm - 1u;
                    // select without dynamic register indexing (g varies per lane)
                    uint32_t accw = (g & 1) ? ao[0] : ae[0];
#pragma unroll
                    for (int i = 1; i < 4; ++i) accw = (g >> 2) == i ? ((g & 1) ? ao[i] : ae[i]) : accw;
                    const uint32_t sum = (g & 2) ? (accw >> 16) : (accw & 0xFFFFu);
                    if (slot < POOL) {
                        pool[slot] = make_uint2(uint32_t(v), uint32_t(g) | (sum << 4));
                    } else {  // pool full: straight to the global buffer
                        const int q = qid[g];
                        const int pos = atomicAdd(a.cand_cnt + q, 1);
                        if (pos < a.capq)
                            a.cand[int64_t(q) * a.capq + pos] = ckey(delt[g] * float(sum), a.list_ids[beg + v]);
                    }
                    ++slot;
                }
            }
            __syncthreads();  // this step's appends are done
            const int filled = *pcnt;
            __syncthreads();  // every thread has read the count before the next appends
            if (filled > POOL - FNT) flush_pool<M>(a, smem, beg);  // uniform
        }
        // the loop's last barrier pair ordered every append before this read,
        // and no thread appends again before the next item's barriers
        if (*pcnt > 0) flush_pool<M>(a, smem, beg);
        __syncthreads();  // the next item overwrites qid, the tables and the pool
    }
}

// ---- seed lists per query
__global__ void seed_count_kernel(const int32_t *__restrict__ probes, int nq, int w,
                                  const int64_t *__restrict__ list_off, int nlist, int kl, int min_seed,
                                  int32_t *__restrict__ ns) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    int64_t have = 0;
    int r = 0;
    for (; r < w; ++r) {
        if (r >= min_seed && have >= kl) break;
        const int l = probes[int64_t(q) * w + r];
        if (l >= 0 && l < nlist) have += list_off[l + 1] - list_off[l];
    }
    ns[q] = r;
}

// ---- list-major grouping of (query, probe) pairs with rank >= rank0[q]
__global__ void pair_count_kernel(const int32_t *__restrict__ probes, int nq, int w,
                                  const int32_t *__restrict__ rank0, int nlist, int *__restrict__ cnt) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(nq) * w) return;
    const int r = int(e % w);
    const int l = probes[e];
    if (r >= rank0[e / w] && l >= 0 && l < nlist) atomicAdd(cnt + l, 1);
}

// exclusive scans of pair counts and of work items per list (one workgroup)
__global__ void pair_scan_kernel(const int *__restrict__ cnt, int nlist, int *__restrict__ pair_off,
                                 int *__restrict__ item_off, int *__restrict__ total_items,
                                 int *__restrict__ fill) {
    __shared__ int part_p[1024], part_i[1024];
    const int tid = threadIdx.x;
    const int per = (nlist + 1023) / 1024;
    int sp = 0, si = 0;
    for (int j = tid * per; j < min(nlist, (tid + 1) * per); ++j) {
        sp += cnt[j];
        si += (cnt[j] + FG - 1) / FG;
    }
    part_p[tid] = sp;
    part_i[tid] = si;
    __syncthreads();
    if (tid == 0) {
        int ap = 0, ai = 0;
        for (int i = 0; i < 1024; ++i) {
            const int tp = part_p[i], ti = part_i[i];
            part_p[i] = ap;
            part_i[i] = ai;
            ap += tp;
            ai += ti;
        }
        pair_off[nlist] = ap;
        item_off[nlist] = ai;
        *total_items = ai;
    }
    __syncthreads();
    int ap = part_p[tid], ai = part_i[tid];
    for (int j = tid * per; j < min(nlist, (tid + 1) * per); ++j) {
        pair_off[j] = ap;
        item_off[j] = ai;
        fill[j] = 0;
        ap += cnt[j];
        ai += (cnt[j] + FG - 1) / FG;
    }
}

__global__ void pair_scatter_kernel(const int32_t *__restrict__ probes, int nq, int w,
                                    const int32_t *__restrict__ rank0, int nlist, const int *__restrict__ pair_off,
                                    int *__restrict__ fill, int32_t *__restrict__ pair_q) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(nq) * w) return;
    const int r = int(e % w);
    const int l = probes[e];
    if (r >= rank0[e / w] && l >= 0 && l < nlist) pair_q[pair_off[l] + atomicAdd(fill + l, 1)] = int32_t(e / w);
}

__global__ void item_expand_kernel(const int *__restrict__ pair_off, const int *__restrict__ item_off, int nlist,
                                   int32_t *__restrict__ item_list, int32_t *__restrict__ item_pair) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlist) return;
    const int n = pair_off[l + 1] - pair_off[l];
    for (int j = 0, it = item_off[l]; j * FG < n; ++j, ++it) {
        item_list[it] = l;
        item_pair[it] = pair_off[l] + j * FG;
    }
}

// ---- seed: tau_q = kl-th exact distance of the seed list (inf if fewer)
__global__ void seed_tau_kernel(const float *__restrict__ sdis, int nq, int kl, float *__restrict__ tau) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq) tau[q] = sdis[int64_t(q) * kl + kl - 1];
}

// ---- per query: the kl best (distance, id) keys of seed + candidates
__global__ __launch_bounds__(256) void fast_select_kernel(const int32_t *__restrict__ sids,
                                                          const float *__restrict__ sdis, int kl,
                                                          const uint64_t *__restrict__ cand,
                                                          const int *__restrict__ cand_cnt, int capq,
                                                          int32_t *__restrict__ short_ids) {
    __shared__ SelectShared sh;
    __shared__ uint64_t buf[1024];
    __shared__ int cnt, eqt;
    const int q = blockIdx.x;
    const int tid = threadIdx.x;
    const int nc = min(cand_cnt[q], capq);
    const uint64_t *C = cand + int64_t(q) * capq;
    const int n = kl + nc;
    auto keyfn = [&](int i) -> uint64_t {
        if (i < kl) {
            const int32_t id = sids[int64_t(q) * kl + i];
            return id < 0 ? kNoKey : ckey(sdis[int64_t(q) * kl + i], id);
        }
        return C[i - kl];
    };
    SelectResult r = block_select<256>(keyfn, n, kl, sh);
    if (tid == 0) {
        cnt = 0;
        eqt = 0;
    }
    __syncthreads();
    for (int i = tid; i < n; i += 256) {
        const uint64_t key = keyfn(i);
        if (key == kNoKey) continue;
        bool take = key < r.thr;
        if (!take && r.eq > 0 && key == r.thr) take = atomicAdd(&eqt, 1) < r.eq;
        if (take) {
            const int pos = atomicAdd(&cnt, 1);
            if (pos < 1024) buf[pos] = key;
        }
    }
    __syncthreads();
    const int c = min(cnt, kl);
    for (int i = tid; i < kl; i += 256) short_ids[int64_t(q) * kl + i] = i < c ? key_id(buf[i]) : -1;
}

inline unsigned nblk(int64_t n, int t) { return unsigned((n + t - 1) / t); }

template <int M>
hipError_t launch_fast_m(const FastArgs &a, int grid, hipStream_t s) {
    constexpr size_t lds = FLay<M>::total;
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&fast_scan_kernel<M>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
        attr = true;
    }
    fast_scan_kernel<M><<<grid, FNT, lds, s>>>(a);
    return hipGetLastError();
}

}  // namespace

int fast_group_size() { return FG; }

hipError_t launch_seed_count(const int32_t *probes, int nq, int w, const int64_t *list_off, int nlist, int kl,
                             int min_seed, int32_t *ns, hipStream_t s) {
    seed_count_kernel<<<nblk(nq, 256), 256, 0, s>>>(probes, nq, w, list_off, nlist, kl, min_seed, ns);
    return hipGetLastError();
}

hipError_t launch_pairs(const int32_t *probes, int nq, int w, const int32_t *rank0, int nlist, int *cnt,
                        int *pair_off, int *item_off, int *total_items, int *fill, int32_t *pair_q,
                        int32_t *item_list, int32_t *item_pair, hipStream_t s) {
    hipError_t e = hipMemsetAsync(cnt, 0, sizeof(int) * nlist, s);
    if (e != hipSuccess) return e;
    const int64_t np = int64_t(nq) * w;
    pair_count_kernel<<<nblk(np, 256), 256, 0, s>>>(probes, nq, w, rank0, nlist, cnt);
    pair_scan_kernel<<<1, 1024, 0, s>>>(cnt, nlist, pair_off, item_off, total_items, fill);
    pair_scatter_kernel<<<nblk(np, 256), 256, 0, s>>>(probes, nq, w, rank0, nlist, pair_off, fill, pair_q);
    item_expand_kernel<<<nblk(nlist, 256), 256, 0, s>>>(pair_off, item_off, nlist, item_list, item_pair);
    return hipGetLastError();
}

hipError_t launch_seed_tau(const float *sdis, int nq, int kl, float *tau, hipStream_t s) {
    seed_tau_kernel<<<nblk(nq, 256), 256, 0, s>>>(sdis, nq, kl, tau);
    return hipGetLastError();
}

hipError_t launch_fast_scan(const FastArgs &a, int m, int grid, hipStream_t s) {
    switch (m) {
        case 8: return launch_fast_m<8>(a, grid, s);
        case 16: return launch_fast_m<16>(a, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_select(const int32_t *sids, const float *sdis, int kl, const uint64_t *cand,
                              const int *cand_cnt, int capq, int nq, int32_t *short_ids, hipStream_t s) {
    if (kl < 1 || kl > 1024) return hipErrorInvalidValue;
    fast_select_kernel<<<nq, 256, 0, s>>>(sids, sdis, kl, cand, cand_cnt, capq, short_ids);
    return hipGetLastError();
}

}  // namespace mivf
