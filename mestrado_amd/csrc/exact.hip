// Batched exact search: per-query w and k, list-major, bit-exact to the plain
// form (and so to the reference arithmetic).
//
// The (list, query) pairs of the whole batch are grouped by list into work
// items of G queries (the same grouping as the reduced-precision form), and
// one persistent kernel takes items from a global counter.  Per item it
// builds the G exact FP32 tables of the reference arithmetic in LDS,
// interleaved as tab[s][code][G], and scans the list once: each lane adds
// the G entries of its vector's code subquantizer by subquantizer (s = 0f;
// s = fl32(s + T[j][c_j]), j ascending, like sumidxtab2), so every distance
// equals the plain form's bit for bit.  A vector is kept for query g when its
// distance is <= tau_g, an upper bound of that query's k-th distance (the
// seed threshold over its nearest lists), so the candidates hold the query's
// exact top-k; exact_select_kernel keeps the k best by (distance, id).
// Queries whose candidates exceed the buffer are flagged; the host reruns
// them with the plain kernels.
#include "select_dev.h"

namespace mivf {
namespace {

constexpr int D = 128;
constexpr int POOLX = 1536;  // staged candidates per workgroup

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const f32x4 lds_f32x4;

template <int M, int G>
struct XLay {
    static constexpr size_t tab = 0;                                // float [M][256][G]
    static constexpr size_t pool = tab + size_t(M) * 256 * G * 4;   // uint2 [POOLX] (scan phase)
    static constexpr size_t R = pool;                               // float [G][D]  (table phase)
    static constexpr size_t rn = R + size_t(G) * D * 4;             // float [G][M]  (table phase)
    static_assert(rn + size_t(G) * M * 4 <= pool + POOLX * 8, "table-phase data must fit in the pool");
    static constexpr size_t prank = pool + POOLX * 8;               // uint16 [POOLX]
    static constexpr size_t lim = prank + POOLX * 2;                // float [16]
    static constexpr size_t qid = lim + 64;                         // int [16]
    static constexpr size_t gcnt = qid + 64;                        // int [16]
    static constexpr size_t gbase = gcnt + 64;                      // int [16]
    static constexpr size_t rec = gbase + 64;                       // FastItem
    static constexpr size_t nrec = rec + sizeof(FastItem);          // FastItem
    static constexpr size_t pcnt = nrec + sizeof(FastItem);         // int [4] rotating append counters
    static constexpr size_t item = pcnt + 16;                       // int
    static constexpr size_t nidx = item + 4;                        // int
    static constexpr size_t hsc = nidx + 12;                        // float [16] histogram scale kExactBins / tau
    static constexpr size_t lhist = hsc + 64;                       // int [G][kExactBins] flush histogram
    static constexpr size_t total = lhist + size_t(G) * kExactBins * 4;
};
// two workgroups of the G = 4 shape share a CU's 160 KiB
static_assert(2 * XLay<16, 4>::total <= 160 * 1024, "two workgroups per CU");

template <int M>
__device__ inline void xload_codes(const uint8_t *lc, int v, uint32_t (&cw)[M / 4]) {
    if constexpr (M == 16) {
        const uint4 c4 = *reinterpret_cast<const uint4 *>(lc + uint32_t(v) * 16u);
        cw[0] = c4.x; cw[1] = c4.y; cw[2] = c4.z; cw[3] = c4.w;
    } else {
        const uint2 c2 = *reinterpret_cast<const uint2 *>(lc + uint32_t(v) * 8u);
        cw[0] = c2.x; cw[1] = c2.y;
    }
}

// Staged candidates -> per-query global buffers as (distance, id) keys.
// Entry: x = list position | g << 28, y = distance bits.
// With running bounds, each flush also adds its distances to the queries'
// histograms (bin b holds [b, b + 1) tau_q / kExactBins) and lowers runthr[q]
// to the upper edge of the first bin whose cumulative count reaches k_q: at
// least k_q kept vectors are then within it, so a vector beyond it cannot be
// among the query's k best (by distance, then id: equal distances stay).
// Counts read back may miss concurrent additions; a smaller count only gives
// a larger, still valid, bound.
template <int M, int G, int NT>
__device__ void xflush(const ExactArgs &a, unsigned char *smem, int64_t beg, int filled) {
    using L = XLay<M, G>;
    const uint2 *pool = reinterpret_cast<const uint2 *>(smem + L::pool);
    uint16_t *prank = reinterpret_cast<uint16_t *>(smem + L::prank);
    int *gcnt = reinterpret_cast<int *>(smem + L::gcnt);
    int *gbase = reinterpret_cast<int *>(smem + L::gbase);
    const int *qid = reinterpret_cast<const int *>(smem + L::qid);
    const float *hsc = reinterpret_cast<const float *>(smem + L::hsc);
    int *lhist = reinterpret_cast<int *>(smem + L::lhist);
    const int tid = threadIdx.x;
    const int n = min(filled, POOLX);
    if (tid < 16) gcnt[tid] = 0;
    if (tid < G * kExactBins) lhist[tid] = 0;
    lds_barrier();
    for (int i = tid; i < n; i += NT) prank[i] = uint16_t(atomicAdd(gcnt + (pool[i].x >> 28), 1));
    lds_barrier();
    if (tid < G && gcnt[tid] > 0) gbase[tid] = atomicAdd(a.cand_cnt + qid[tid], gcnt[tid]);
    lds_barrier();
    for (int i = tid; i < n; i += NT) {
        const uint2 e = pool[i];
        const int g = int(e.x >> 28);
        const int pos = gbase[g] + int(prank[i]);
        const float dist = __builtin_bit_cast(float, e.y);
        if (pos < a.capq) {
            a.cand[int64_t(qid[g]) * a.capq + pos] = ckey(dist, a.list_ids[beg + (e.x & 0x0FFFFFFFu)]);
            if (hsc[g] > 0.0f) atomicAdd(lhist + g * kExactBins + min(kExactBins - 1, int(fmaxf(dist, 0.0f) * hsc[g])), 1);
        }
    }
    lds_barrier();  // the pool and the flush histograms are complete
    if (a.runthr) {
        if (tid < G * kExactBins) {
            const int c = lhist[tid];
            if (c > 0) atomicAdd(a.rhist + int64_t(qid[tid / kExactBins]) * kExactBins + (tid % kExactBins), c);
        }
        const int lane = tid & 63;
        for (int g = tid >> 6; g < G; g += NT / 64) {  // one wavefront per slot
            if (!(hsc[g] > 0.0f) || gcnt[g] == 0) continue;  // wave-uniform
            const int q = qid[g];
            const int kq = a.kq ? a.kq[q] : a.kuni;
            int c = lane < kExactBins ? __hip_atomic_load(a.rhist + int64_t(q) * kExactBins + lane, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT) : 0;
#pragma unroll
            for (int off = 1; off < kExactBins; off <<= 1) {
                const int o = __shfl_up(c, off, 64);
                if (lane >= off) c += o;
            }
            const uint64_t reach = __ballot(lane < kExactBins && c >= kq);
            if (reach != 0 && lane == 0) {
                const int b = __ffsll((unsigned long long)reach) - 1;
                const float R = (float(b + 1) + 0.01f) / hsc[g];
                atomicMin(reinterpret_cast<int *>(a.runthr) + q, __float_as_int(R));
            }
        }
    }
    lds_barrier();  // the pool is free again
}

// The G exact tables of one item (reference arithmetic, as the plain scan's
// table phase): r = fl(q - c_l); |r_s|^2 = sum_i fl(r_i^2) in order;
// dot = sum_i fl(r_i p_i) in order; T = fl(fl(-2 dot) + fl(|p|^2_f64 + |r_s|^2))
// (-2 dot is exact, so fma(-2, dot, cc) rounds once, identically).
// Each thread owns one subquantizer s and E codes of it (c0, c0 + CPS, ...),
// so the residual slices r_s of GC queries are loaded into registers once and
// reused for all E codes; consecutive lanes take consecutive codes (coalesced
// centroid loads, conflict-free 16-byte table stores).  The dot starts from
// the first product: an all-zero dot cannot change an entry, since the norm
// term is never -0.
template <int M, int G, int NT>
__device__ void xtables(const ExactArgs &a, unsigned char *smem, int l) {
    constexpr int DS = D / M;
    constexpr int E = M * 256 / NT;                    // codes per thread
    constexpr int CPS = 256 / E;                       // threads per subquantizer
    constexpr int GC = (G * DS <= 64) ? G : 64 / DS;   // queries per register-resident chunk
    static_assert(E * NT == M * 256 && CPS * M == NT, "one subquantizer per thread");
    static_assert(G % GC == 0 && GC % 4 == 0, "query chunks of whole float4s");
    using L = XLay<M, G>;
    float *R = reinterpret_cast<float *>(smem + L::R);
    float *rn = reinterpret_cast<float *>(smem + L::rn);
    const int *qid = reinterpret_cast<const int *>(smem + L::qid);
    const int tid = threadIdx.x;
    for (int e = tid; e < G * D; e += NT) {
        const int g = e / D, t = e % D;
        const int q = qid[g];
        R[e] = q >= 0 ? a.queries[int64_t(q) * D + t] - a.coarse[int64_t(l) * D + t] : 0.0f;
    }
    __syncthreads();
    if (tid < G * M) {
        const int g = tid / M, s = tid % M;
        float acc = 0.0f;
#pragma unroll
        for (int i = 0; i < DS; ++i) {
            const float v = R[g * D + s * DS + i];
            const float sq = v * v;
            acc = acc + sq;
        }
        rn[g * M + s] = acc;
    }
    __syncthreads();
    const int s = tid / CPS, c0 = tid % CPS;
#pragma unroll
    for (int g0 = 0; g0 < G; g0 += GC) {
        float r[GC][DS];
        double rn64[GC];
#pragma unroll
        for (int g = 0; g < GC; ++g) {
#pragma unroll
            for (int i = 0; i < DS; ++i) r[g][i] = R[(g0 + g) * D + s * DS + i];
            rn64[g] = double(rn[(g0 + g) * M + s]);
        }
        // all E codes unrolled: their centroid rows load together (a
        // prefetch of the next code's row into copies gains nothing: the
        // copies cost VALU instructions, which bound this phase)
#pragma unroll
        for (int j = 0; j < E; ++j) {
            const int c = c0 + CPS * j;
            float vals[GC];
            if (c < a.ks) {
                float pr[DS];
                const float4 *src = reinterpret_cast<const float4 *>(a.pq + (int64_t(s) * a.ks + c) * DS);
#pragma unroll
                for (int i = 0; i < DS / 4; ++i) {
                    const float4 v = src[i];
                    pr[4 * i] = v.x;
                    pr[4 * i + 1] = v.y;
                    pr[4 * i + 2] = v.z;
                    pr[4 * i + 3] = v.w;
                }
                const double pnv = a.pq_norm[int64_t(s) * a.ks + c];
#pragma unroll
                for (int g = 0; g < GC; ++g) {
                    float dot = r[g][0] * pr[0];
#pragma unroll
                    for (int i = 1; i < DS; ++i) {
                        const float prod = r[g][i] * pr[i];
                        dot = dot + prod;
                    }
                    const float cc = float(pnv + rn64[g]);
                    vals[g] = __builtin_fmaf(-2.0f, dot, cc);
                }
            } else {
#pragma unroll
                for (int g = 0; g < GC; ++g) vals[g] = 0.0f;  // codes are < ks: never indexed
            }
            float4 *dst = reinterpret_cast<float4 *>(smem + L::tab + (size_t(s * 256 + c) * G + g0) * 4);
#pragma unroll
            for (int g4 = 0; g4 < GC / 4; ++g4)
                dst[g4] = make_float4(vals[4 * g4], vals[4 * g4 + 1], vals[4 * g4 + 2], vals[4 * g4 + 3]);
        }
    }
    __syncthreads();  // tables ready; the pool may now overwrite R and rn
}

// Waves per SIMD the LDS allows: two workgroups per CU when two tables fit
// in the 160 KiB, else one.
template <int M, int G, int NT>
constexpr int xwaves() {
    return (2 * XLay<M, G>::total <= 160 * 1024 ? 2 : 1) * NT / 256;
}

template <int M, int G, int NT>
__global__ __launch_bounds__(NT, (xwaves<M, G, NT>())) void exact_scan_kernel(ExactArgs a) {
    using L = XLay<M, G>;
    constexpr int RW = int(sizeof(FastItem) / 4);
    static_assert(G % 4 == 0 && G <= 16, "G must be 4, 8, 12 or 16");
    extern __shared__ __align__(16) unsigned char smem[];
    float *lim = reinterpret_cast<float *>(smem + L::lim);
    float *hsc = reinterpret_cast<float *>(smem + L::hsc);
    int *qid = reinterpret_cast<int *>(smem + L::qid);
    int *pcnt = reinterpret_cast<int *>(smem + L::pcnt);
    int *item_sh = reinterpret_cast<int *>(smem + L::item);
    int *nidx_sh = reinterpret_cast<int *>(smem + L::nidx);
    FastItem *rec = reinterpret_cast<FastItem *>(smem + L::rec);
    FastItem *nrec = reinterpret_cast<FastItem *>(smem + L::nrec);
    uint2 *pool = reinterpret_cast<uint2 *>(smem + L::pool);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    if (uint32_t(uintptr_t((__attribute__((address_space(3))) unsigned char *)(smem))) != 0) {
        if (tid == 0) atomicOr(a.err, 4);
        return;
    }
    const int total = *a.total_items;
    if (tid < 4) pcnt[tid] = 0;
    if (tid == 0) *item_sh = atomicAdd(a.next_item, 1);
    __syncthreads();
    if (*item_sh < total && tid < RW)
        reinterpret_cast<uint32_t *>(rec)[tid] = reinterpret_cast<const uint32_t *>(a.items + *item_sh)[tid];
    __syncthreads();

    for (;;) {
        const int item = *item_sh;
        if (item >= total) break;  // uniform
        const int64_t beg = rec->beg;
        const int len = rec->len;
        const int l = rec->l;
        if (tid < 16) {
            const int q = tid < G ? rec->q[tid] : -1;
            qid[tid] = q;
            const float tau = q >= 0 ? (a.tau ? a.tau[q] : rec->tau[tid]) : -__builtin_huge_valf();
            // the running bound R_q (>= k_q kept vectors within it; +inf until then)
            const float R = (q >= 0 && a.runthr) ? __hip_atomic_load(a.runthr + q, __ATOMIC_RELAXED,
                                                                     __HIP_MEMORY_SCOPE_AGENT) : __builtin_huge_valf();
            lim[tid] = fminf(tau, R);
            hsc[tid] = (q >= 0 && a.runthr && tau > 0.0f && tau < 3.0e38f) ? float(kExactBins) / tau : 0.0f;
        }
        if (tid < 3) pcnt[tid] = 0;
        __syncthreads();
        xtables<M, G, NT>(a, smem, l);

        // the next item's index and record are fetched during the scan
        int nxt = 0;
        if (tid == 0) nxt = atomicAdd(a.next_item, 1);
        int stage = 0;
        uint32_t rv = 0u;
        auto advance = [&]() {
            if (stage == 0) {
                if (wave == 0) {
                    const int nb = __shfl(nxt, 0);
                    if (lane < RW && nb < total) rv = reinterpret_cast<const uint32_t *>(a.items + nb)[lane];
                    if (lane == 0) *nidx_sh = nb;
                }
            } else if (stage == 1) {
                if (wave == 0 && lane < RW) reinterpret_cast<uint32_t *>(nrec)[lane] = rv;
            }
            ++stage;
        };

        float lm[G];
#pragma unroll
        for (int g = 0; g < G; ++g) lm[g] = lim[g];
        const uint8_t *lc = a.codes + beg * M;
        uint32_t cn[M / 4];
        xload_codes<M>(lc, min(tid, max(len - 1, 0)), cn);
        int fill = 0;  // staged candidates before this step (uniform)
        int rot = 0;   // append counter of this step
        for (int base = 0; base < len; base += NT) {
            if (tid == 0) pcnt[rot == 2 ? 0 : rot + 1] = 0;
            const int v = base + tid;
            uint32_t cw[M / 4];
#pragma unroll
            for (int i = 0; i < M / 4; ++i) cw[i] = cn[i];
            if (base + NT < len) xload_codes<M>(lc, min(v + NT, len - 1), cn);
            float acc[G];
#pragma unroll
            for (int s = 0; s < M; ++s) {
                const uint32_t code = (cw[s >> 2] >> (8 * (s & 3))) & 255u;
                const uint32_t row = uint32_t(L::tab) + (uint32_t(s) * 256u + code) * uint32_t(G * 4);
#pragma unroll
                for (int g4 = 0; g4 < G / 4; ++g4) {
                    const f32x4 x = *reinterpret_cast<lds_f32x4 *>(uintptr_t(row + 16u * g4));
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[4 * g4 + j] = s == 0 ? x[j] : acc[4 * g4 + j] + x[j];  // s = 0f; s = fl(s + T)
                }
            }
            uint32_t pm = 0u;
#pragma unroll
            for (int g = 0; g < G; ++g) pm |= uint32_t(acc[g] <= lm[g]) << g;
            if (v >= len) pm = 0u;
            if (__ballot(pm != 0u) != 0) {  // wave-uniform
                // wave prefix sum of the per-lane pass counts (<= G: the bit
                // planes of G)
                constexpr int NPL = G <= 1 ? 1 : 32 - __builtin_clz(unsigned(G));
                const uint32_t n = uint32_t(__popc(pm));
                uint32_t pre = 0u, tot = 0u;
#pragma unroll
                for (int b = 0; b < NPL; ++b) {
                    const uint64_t bm = __ballot((n >> b) & 1u);
                    pre += uint32_t(__builtin_amdgcn_mbcnt_hi(uint32_t(bm >> 32),
                                                              __builtin_amdgcn_mbcnt_lo(uint32_t(bm), 0u))) << b;
                    tot += uint32_t(__popcll(bm)) << b;
                }
                int slot = 0;
                if (lane == 0) slot = atomicAdd(pcnt + rot, int(tot));
                slot = fill + __builtin_amdgcn_readfirstlane(slot) + int(pre);
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    if (!((pm >> g) & 1u)) continue;
                    if (slot < POOLX) {
                        pool[slot] = make_uint2(uint32_t(v) | (uint32_t(g) << 28), __builtin_bit_cast(uint32_t, acc[g]));
                    } else {  // pool full: straight to the global buffer
                        const int q = qid[g];
                        const int pos = atomicAdd(a.cand_cnt + q, 1);
                        if (pos < a.capq) a.cand[int64_t(q) * a.capq + pos] = ckey(acc[g], a.list_ids[beg + v]);
                    }
                    ++slot;
                }
            }
            // Three append counters rotate over the steps: counter rot takes
            // this step's appends, counter rot + 1 (cleared at this step's
            // start) the next step's, and counter rot + 2, read by every
            // thread before this step's barrier, is free.  One LDS-only
            // barrier per step then orders the appends before the reads.
            lds_barrier();  // this step's appends are done
            fill += pcnt[rot];
            rot = rot == 2 ? 0 : rot + 1;
            if (fill > POOLX / 2) {  // uniform
                xflush<M, G, NT>(a, smem, beg, fill);
                fill = 0;
            }
            if (stage < 2) advance();
        }
        if (fill > 0) xflush<M, G, NT>(a, smem, beg, fill);
        while (stage < 2) {
            __syncthreads();  // orders the previous stage's LDS writes before this stage
            advance();
        }
        __syncthreads();  // rec, qid and the pool are free; nrec is complete
        if (tid < RW) reinterpret_cast<uint32_t *>(rec)[tid] = reinterpret_cast<const uint32_t *>(nrec)[tid];
        if (tid == 0) *item_sh = *nidx_sh;
        __syncthreads();
    }
}

// Test hook: the G exact tables of queries 0..n-1 against list l, copied out.
template <int M, int G, int NT>
__global__ __launch_bounds__(NT) void exact_tables_kernel(ExactArgs a, int l, int n, float *out) {
    using L = XLay<M, G>;
    extern __shared__ __align__(16) unsigned char smem[];
    int *qid = reinterpret_cast<int *>(smem + L::qid);
    const int tid = threadIdx.x;
    if (tid < 16) qid[tid] = tid < n && tid < G ? tid : -1;
    __syncthreads();
    xtables<M, G, NT>(a, smem, l);
    for (int i = tid; i < M * 256 * G; i += NT) {
        const uint32_t addr = uint32_t(L::tab) + uint32_t(i) * 4u;
        out[i] = *reinterpret_cast<__attribute__((address_space(3))) const float *>(uintptr_t(addr));
    }
}

// Test hook: ADC of every vector of list l for queries 0..n-1 (out [len][G]),
// with the scan loop's table reads.
template <int M, int G, int NT>
__global__ __launch_bounds__(NT) void exact_adc_kernel(ExactArgs a, int l, int n, int64_t beg, int len, float *out) {
    using L = XLay<M, G>;
    extern __shared__ __align__(16) unsigned char smem[];
    int *qid = reinterpret_cast<int *>(smem + L::qid);
    const int tid = threadIdx.x;
    if (tid < 16) qid[tid] = tid < n && tid < G ? tid : -1;
    __syncthreads();
    xtables<M, G, NT>(a, smem, l);
    const uint8_t *lc = a.codes + beg * M;
    for (int v = tid; v < len; v += NT) {
        uint32_t cw[M / 4];
        xload_codes<M>(lc, v, cw);
        float acc[G];
#pragma unroll
        for (int g = 0; g < G; ++g) acc[g] = 0.0f;
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const uint32_t code = (cw[s >> 2] >> (8 * (s & 3))) & 255u;
            const uint32_t row = uint32_t(L::tab) + (uint32_t(s) * 256u + code) * uint32_t(G * 4);
#pragma unroll
            for (int g4 = 0; g4 < G / 4; ++g4) {
                const f32x4 x = *reinterpret_cast<lds_f32x4 *>(uintptr_t(row + 16u * g4));
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[4 * g4 + j] = acc[4 * g4 + j] + x[j];
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) out[int64_t(v) * G + g] = acc[g];
    }
}

__global__ void mask_probes_kernel(int32_t *__restrict__ probes, int nq, int w, const int32_t *__restrict__ wq) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(nq) * w) return;
    if (int(e % w) >= wq[e / w]) probes[e] = -1;
}

// Per query: the k_q best (distance, id) of its candidates, sorted; rows are
// kstride wide, padded with (-1, +inf).  Overflowed queries are flagged.
__global__ __launch_bounds__(256) void exact_select_kernel(const uint64_t *__restrict__ cand,
                                                           const int *__restrict__ cand_cnt, int capq,
                                                           const int32_t *__restrict__ kq, int kuni, int kstride,
                                                           int32_t *__restrict__ out_ids, float *__restrict__ out_dis,
                                                           int *__restrict__ overflow) {
    __shared__ SelectShared sh;
    __shared__ uint64_t buf[1024];
    __shared__ int cnt, eqt;
    const int q = blockIdx.x;
    const int tid = threadIdx.x;
    const int n = cand_cnt[q];
    const int k = kq ? kq[q] : kuni;
    if (n > capq) {  // uniform: the host reruns this query
        if (tid == 0) {
            overflow[q] = 1;
            atomicAdd(overflow + gridDim.x, 1);
        }
        return;
    }
    if (tid == 0) overflow[q] = 0;
    const uint64_t *C = cand + int64_t(q) * capq;
    auto keyfn = [&](int i) -> uint64_t { return C[i]; };
    const SelectResult r = block_select<256>(keyfn, n, k, sh);
    if (tid == 0) {
        cnt = 0;
        eqt = 0;
    }
    __syncthreads();
    for (int i = tid; i < n; i += 256) {
        const uint64_t key = C[i];
        bool take = key < r.thr;
        if (!take && r.eq > 0 && key == r.thr) take = atomicAdd(&eqt, 1) < r.eq;
        if (take) {
            const int pos = atomicAdd(&cnt, 1);
            if (pos < 1024) buf[pos] = key;
        }
    }
    __syncthreads();
    const int c = min(cnt, 1024);
    const int np = next_pow2(c > 0 ? c : 1);
    for (int i = c + tid; i < np; i += 256) buf[i] = kNoKey;
    __syncthreads();
    block_bitonic_sort<256>(buf, np);
    for (int i = tid; i < kstride; i += 256) {
        const bool ok = i < c && i < k;
        out_ids[int64_t(q) * kstride + i] = ok ? key_id(buf[i]) : -1;
        out_dis[int64_t(q) * kstride + i] = ok ? key_dist(buf[i]) : __builtin_huge_valf();
    }
}

template <int M, int G, int NT>
hipError_t launch_x(const ExactArgs &a, int grid, hipStream_t s) {
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&exact_scan_kernel<M, G, NT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(XLay<M, G>::total));
        if (e != hipSuccess) return e;
        attr = true;
    }
    exact_scan_kernel<M, G, NT><<<grid, NT, XLay<M, G>::total, s>>>(a);
    return hipGetLastError();
}

}  // namespace

size_t exact_lds_bytes(int m, int g) {
    switch (m * 100 + g) {
        case 1604: return XLay<16, 4>::total;
        case 1608: return XLay<16, 8>::total;
        case 804: return XLay<8, 4>::total;
        case 808: return XLay<8, 8>::total;
        case 816: return XLay<8, 16>::total;
        default: return 0;
    }
}

hipError_t launch_exact_scan(const ExactArgs &a, int m, int g, int nt, int grid, hipStream_t s) {
    switch ((m * 100 + g) * 10000 + nt) {
        case 16040256: return launch_x<16, 4, 256>(a, grid, s);
        case 16040512: return launch_x<16, 4, 512>(a, grid, s);
        case 8040512: return launch_x<8, 4, 512>(a, grid, s);
        case 16080256: return launch_x<16, 8, 256>(a, grid, s);
        case 16080512: return launch_x<16, 8, 512>(a, grid, s);
        case 8040256: return launch_x<8, 4, 256>(a, grid, s);
        case 8080256: return launch_x<8, 8, 256>(a, grid, s);
        case 8080512: return launch_x<8, 8, 512>(a, grid, s);
        case 8160512: return launch_x<8, 16, 512>(a, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_exact_tables(const ExactArgs &a, int m, int l, int n, float *out, hipStream_t s) {
    if (m == 16) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&exact_tables_kernel<16, 8, 512>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(XLay<16, 8>::total));
        if (e != hipSuccess) return e;
        exact_tables_kernel<16, 8, 512><<<1, 512, XLay<16, 8>::total, s>>>(a, l, n, out);
    } else if (m == 8) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&exact_tables_kernel<8, 8, 512>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(XLay<8, 8>::total));
        if (e != hipSuccess) return e;
        exact_tables_kernel<8, 8, 512><<<1, 512, XLay<8, 8>::total, s>>>(a, l, n, out);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_exact_adc(const ExactArgs &a, int m, int l, int n, int64_t beg, int len, float *out, hipStream_t s) {
    if (m != 16) return hipErrorInvalidValue;
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&exact_adc_kernel<16, 8, 512>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, int(XLay<16, 8>::total));
    if (e != hipSuccess) return e;
    exact_adc_kernel<16, 8, 512><<<1, 512, XLay<16, 8>::total, s>>>(a, l, n, beg, len, out);
    return hipGetLastError();
}

hipError_t launch_mask_probes(int32_t *probes, int nq, int w, const int32_t *wq, hipStream_t s) {
    const int64_t n = int64_t(nq) * w;
    if (n <= 0) return hipSuccess;
    mask_probes_kernel<<<unsigned((n + 255) / 256), 256, 0, s>>>(probes, nq, w, wq);
    return hipGetLastError();
}

hipError_t launch_exact_select(const uint64_t *cand, const int *cand_cnt, int capq, const int32_t *kq, int kuni,
                               int kstride, int nq, int32_t *out_ids, float *out_dis, int *overflow, hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    if (kstride < 1 || kstride > 1024 || kuni > 1024) return hipErrorInvalidValue;
    exact_select_kernel<<<unsigned(nq), 256, 0, s>>>(cand, cand_cnt, capq, kq, kuni, kstride, out_ids, out_dis,
                                                     overflow);
    return hipGetLastError();
}

}  // namespace mivf
