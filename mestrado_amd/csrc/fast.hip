// Batched reduced-precision IVF-PQ search (approximate scan + exact rerank).
//
// Work is list-major: a work item is (list l, up to FG = 16 queries that probe
// l).  Its 16 distance tables are quantized to 8 bits and interleaved in LDS
// as t8[s][code][query], so one ds_read_b128 gathers the entries of all 16
// queries for a vector's code (the code belongs to the vector, not to the
// query), and the list's codes are read from HBM once per 16 queries.
//
// Tables.  With r = q - c_l, sigma_s = max_c |p_sc| (1 if 0), bound_g =
// max_s (|r_s| + sigma_s)^2 * 1.0001 and invd_g = 255 / bound_g, the entry
// invd_g * T = invd_g * (|r_s|^2 + |p_sc|^2 - 2 r_s.p_sc) is one dot product
//   a = [-2 invd sigma r_s, invd |r_s|^2, invd sigma^2] . b = [p_sc / sigma, 1, |p_sc|^2 / sigma^2]
// computed on the matrix cores in FP16 with FP32 accumulation
// (v_mfma_f32_16x16x16_f16: 16 query slots x 16 codes per instruction).
// Every |a_k b_k| sums to <= invd (|r_s| + |p_sc|)^2 <= 255 and FP16 rounds
// each operand to 2^-11, so the entry is within 255 * 2^-10 < 0.25 of
// invd * T; rounding to 8 bits adds 0.5, so each table entry is within 0.75
// quanta of invd * T and a vector's 8-bit sum within 0.75 m quanta of
// invd * ADC.  B operands are prepared once per index (fast_prep_kernel).
//
// Sums.  The entries are stored as signed bytes t - 128 and summed over
// subquantizers on the matrix cores (v_mfma_i32_16x16x64_i8 with a one-hot B
// operand), exact in int32; sum = D + 128 m.
//
// Candidate filter: a per-query threshold tau_q >= the kl-th smallest ADC
// distance over a sample of the query's nearest list (seed_tau_kernel).  A
// vector passes if its sum <= tau_q * invd + slack (option fast_seed_slack,
// default 2 quanta): its 8-bit distance is within slack quanta of tau_q.
// slack = 3m/4 + 2 keeps every vector whose ADC distance is <= tau_q (the +2
// covers FP32 differences from the reference table arithmetic); since tau_q
// is the kl-th distance of a sample, it sits above the kl-th of all probed
// vectors, and the default drops only vectors near it, beyond the shortlist
// in practice (recall unchanged on the 10M x 128 workload; 23 % fewer
// candidates).  The candidates are those passing the filter, whatever the
// schedule.  Passing (approximate distance, list position) keys are staged
// in LDS and flushed to per-query global buffers; fast_select keeps the kl
// best and the exact FP32 rerank orders them by (distance, id).
#include "select_dev.h"
#include "rerank_dev.h"

namespace mivf {

namespace {

constexpr int D = 128;
constexpr int FG = 16;    // queries per work item
constexpr int FNT = 512;  // threads per workgroup
constexpr int POOL = 1280;
constexpr int NBH = 64;   // bins of the running-threshold histograms

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const u32x4 lds_u32x4;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const i32x4 lds_i32x4;

// K of the table dot products, padded to whole 16-wide MFMA steps
template <int M>
struct FK {
    static constexpr int DS = D / M;
    static constexpr int KP = (DS + 2 + 15) / 16 * 16;
    static constexpr int NMF = KP / 16;
};

__device__ inline float sigma_of(float pmax) { return pmax > 0.0f ? pmax : 1.0f; }

template <int M>
struct FLay {
    static constexpr int KP = FK<M>::KP;
    static constexpr size_t t8 = 0;                            // uint8 [M][256][FG]
    static constexpr size_t pool = t8 + size_t(M) * 256 * FG;  // uint2 [POOL] (scan phase)
    static constexpr size_t ao = pool;                         // fp16 [M][FG][KP] A operands (table phase)
    static_assert(size_t(M) * FG * KP * 2 <= POOL * 8, "A operands must fit in the pool");
    static constexpr size_t delt = pool + POOL * 8;            // float [FG]
    static constexpr size_t pmx = delt + FG * 4;               // float [M]
    static constexpr size_t stmp = pmx + M * 4;                // uint32 [FG] per-slot limits
    static constexpr size_t qid = stmp + FG * 4;               // int [FG] query of each slot (-1 empty)
    static constexpr size_t gcnt = qid + FG * 4;               // int [FG]
    static constexpr size_t gbase = gcnt + FG * 4;             // int [FG]
    static constexpr size_t hsc = gbase + FG * 4;              // float [FG] histogram scale NBH / tau (0: none)
    static constexpr size_t lhist = hsc + FG * 4;              // int [FG][NBH] flush histogram
    static constexpr size_t rec = lhist + FG * NBH * 4;        // FastItem: the current item
    static constexpr size_t nrec = rec + sizeof(FastItem);     // FastItem: the next item
    static constexpr size_t pcnt = nrec + sizeof(FastItem);    // int [4] rotating append counters
    static constexpr size_t item = pcnt + 16;                  // int current item index
    static constexpr size_t nidx = item + 4;                   // int next item index
    static constexpr size_t total = nidx + 8;
};

// Move the filled staged candidates to the per-query global buffers (whole
// workgroup, uniform call).  Entry: x = list position, y = g | sum << 4,
// and the rank within its query slot goes into bits 16..27 of y.
template <int M>
__device__ void flush_pool(const FastArgs &a, unsigned char *smem, int64_t beg, int filled) {
    using L = FLay<M>;
    uint2 *pool = reinterpret_cast<uint2 *>(smem + L::pool);
    int *gcnt = reinterpret_cast<int *>(smem + L::gcnt);
    int *gbase = reinterpret_cast<int *>(smem + L::gbase);
    const int *qid = reinterpret_cast<const int *>(smem + L::qid);
    const float *delt = reinterpret_cast<const float *>(smem + L::delt);
    const float *hsc = reinterpret_cast<const float *>(smem + L::hsc);
    int *lhist = reinterpret_cast<int *>(smem + L::lhist);
    const int tid = threadIdx.x;
    const int n = min(filled, POOL);
    if (tid < FG) gcnt[tid] = 0;
    for (int i = tid; i < FG * NBH; i += FNT) lhist[i] = 0;
    lds_barrier();
    for (int i = tid; i < n; i += FNT) {
        const uint32_t y = pool[i].y;
        const int r = atomicAdd(gcnt + (y & 15u), 1);
        pool[i].y = y | (uint32_t(r) << 16);
    }
    lds_barrier();
    if (tid < FG && gcnt[tid] > 0) gbase[tid] = atomicAdd(a.cand_cnt + qid[tid], gcnt[tid]);
    lds_barrier();
    for (int i = tid; i < n; i += FNT) {
        const uint2 e = pool[i];
        const int g = int(e.y & 15u);
        const uint32_t sum = (e.y >> 4) & 0xFFFu;
        const int pos = gbase[g] + int(e.y >> 16);
        const float ad = delt[g] * float(sum);
        if (pos < a.capq) {
            a.cand[int64_t(qid[g]) * a.capq + pos] = ckey(ad, int32_t(beg + e.x));
            if (hsc[g] > 0.0f) atomicAdd(lhist + g * NBH + min(NBH - 1, int(ad * hsc[g])), 1);
        }
    }
    lds_barrier();  // the pool is free again
    if (a.runthr) {
        // Running bound per query: add this flush's counts to the query's
        // global histogram of candidate distances (bin b holds distances in
        // [b, b + 1) tau / NBH).  Nothing is read back here: the work items
        // that start later derive their bounds from the histograms
        // (item_bounds), so the flush never waits on these atomics.
        for (int i = tid; i < FG * NBH; i += FNT) {
            const int c = lhist[i];
            if (c > 0) atomicAdd(a.rhist + int64_t(qid[i / NBH]) * NBH + (i % NBH), c);
        }
    }
}

// Slots per wavefront of the item bounds: each wavefront reads the
// histograms of FG / (FNT / 64) query slots.
constexpr int HSW = FG / (FNT / 64);

// Running bounds of a work item's queries: the loads go out at the item's
// start (load_item_hist) and are used once its tables are built
// (item_bounds), so their latency hides behind the table phase.  In a
// query's histogram, the first bin whose cumulative count reaches kl bounds
// the kl-th smallest candidate distance emitted so far: R = its upper edge
// (+1 %), and a vector with dlt * sum above R (with rounding room) cannot
// enter the shortlist.  Counts may miss concurrent additions; a smaller
// count only gives a larger (still valid) bound.
__device__ inline void load_item_hist(const FastArgs &a, const int *qid, int (&hc)[HSW]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < HSW; ++j) {
        const int q = qid[wave + j * (FNT / 64)];
        hc[j] = (a.runthr && q >= 0 && lane < NBH - 1)
                    ? __hip_atomic_load(a.rhist + int64_t(q) * NBH + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                    : 0;
    }
}

template <int M>
__device__ inline void item_bounds(const FastArgs &a, unsigned char *smem, int (&hc)[HSW]) {
    using L = FLay<M>;
    const int *qid = reinterpret_cast<const int *>(smem + L::qid);
    const float *hsc = reinterpret_cast<const float *>(smem + L::hsc);
    const float *delt = reinterpret_cast<const float *>(smem + L::delt);
    uint32_t *stmp = reinterpret_cast<uint32_t *>(smem + L::stmp);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < HSW; ++j) {
        const int g = wave + j * (FNT / 64);
        if (qid[g] < 0 || !(hsc[g] > 0.0f)) continue;  // wave-uniform
        int c = hc[j];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(c, off, 64);
            if (lane >= off) c += o;
        }
        const uint64_t reach = __ballot(lane < NBH - 1 && c >= a.kl);
        if (reach != 0 && lane == 0) {
            const int b = __ffsll((unsigned long long)reach) - 1;
            const float R = (float(b + 1) + 0.01f) / hsc[g];
            const uint32_t lr = uint32_t(fminf(R / delt[g] * 1.0002f + 1.0f, 65534.0f)) + 1u;
            stmp[g] = min(stmp[g], lr);
        }
    }
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

// The M/4 code bytes of subquantizers kb*M/4 .. of vector v (one dword for
// m = 16, a u16 for m = 8), low byte first.
template <int M>
__device__ inline uint32_t load_tcodes(const uint8_t *lc, int v, int kb) {
    if constexpr (M == 16) return *reinterpret_cast<const uint32_t *>(lc + uint32_t(v) * 16u + uint32_t(kb) * 4u);
    else return *reinterpret_cast<const uint16_t *>(lc + uint32_t(v) * 8u + uint32_t(kb) * 2u);
}

// Inputs of one work item's A side: thread (g, s) = (tid / M, tid % M),
// tid < FG * M, holds the ds query and centroid values of its sub-vector.
template <int M>
struct ItemIn {
    static constexpr int DS = D / M;
    float qx[DS], cx[DS];
    __device__ void load(const FastArgs &a, int l, int q, int s) {
#pragma unroll
        for (int t = 0; t < DS; t += 4) {
            const float4 c4 = *reinterpret_cast<const float4 *>(a.coarse + int64_t(l) * D + s * DS + t);
            cx[t] = c4.x; cx[t + 1] = c4.y; cx[t + 2] = c4.z; cx[t + 3] = c4.w;
            float4 q4 = c4;  // empty slot: r = 0
            if (q >= 0) q4 = *reinterpret_cast<const float4 *>(a.queries + int64_t(q) * D + s * DS + t);
            qx[t] = q4.x; qx[t + 1] = q4.y; qx[t + 2] = q4.z; qx[t + 3] = q4.w;
        }
    }
};

// The work item's A operands, scales and limits (from the preloaded
// inputs), then its 8-bit tables in LDS t8.  Uniform call; ends after a
// barrier with the tables complete.
template <int M>
__device__ void build_item_tables(const FastArgs &a, unsigned char *smem, const ItemIn<M> &in,
                                  const float *taus) {
    constexpr int DS = D / M;
    constexpr int KP = FK<M>::KP;
    constexpr int NMF = FK<M>::NMF;
    using L = FLay<M>;
    _Float16 *ao = reinterpret_cast<_Float16 *>(smem + L::ao);
    float *delt = reinterpret_cast<float *>(smem + L::delt);
    const float *pmx = reinterpret_cast<const float *>(smem + L::pmx);
    uint32_t *stmp = reinterpret_cast<uint32_t *>(smem + L::stmp);
    float *hsc = reinterpret_cast<float *>(smem + L::hsc);
    const int *qid = reinterpret_cast<const int *>(smem + L::qid);
    const int tid = threadIdx.x;

    if (tid < FG * M) {
        const int g = tid / M, s = tid % M;
        float r[DS];
        float A = 0.0f;
#pragma unroll
        for (int t = 0; t < DS; ++t) {
            r[t] = in.qx[t] - in.cx[t];
            A = __builtin_fmaf(r[t], r[t], A);
        }
        const float sg = sigma_of(pmx[s]);
        const float b = __builtin_sqrtf(A) + pmx[s];
        float bound = b * b;
#pragma unroll
        for (int off = M / 2; off > 0; off >>= 1) bound = fmaxf(bound, __shfl_xor(bound, off, M));
        bound = fmaxf(bound * 1.0001f, 1e-20f);
        const float invd = 255.0f / bound;
        _Float16 *row = ao + (size_t(s) * FG + g) * KP;
        const float m2 = -2.0f * invd * sg;
#pragma unroll
        for (int t = 0; t < DS; ++t) row[t] = _Float16(m2 * r[t]);
        row[DS] = _Float16(invd * A);
        row[DS + 1] = _Float16(invd * sg * sg);
#pragma unroll
        for (int t = DS + 2; t < KP; ++t) row[t] = _Float16(0.0f);
        if (s == 0) {
            const float dlt = bound / 255.0f;
            delt[g] = dlt;
            const float tau = taus ? taus[g] : (qid[g] >= 0 ? a.tau[qid[g]] : -1.0f);
            const bool seeded = tau < 3.0e38f;
            // stored limit = thr + 1 (a sum passes iff sum < limit)
            uint32_t lim;
            if (qid[g] < 0) lim = 0u;                     // empty slot: nothing passes
            else if (!seeded) lim = 0xFFFFu;              // no seed bound: keep everything
            else lim = uint32_t(fminf(tau * invd + a.slack, 65534.0f)) + 1u;
            stmp[g] = lim;  // lowered by the running bound after the tables (item_bounds)
            hsc[g] = (qid[g] >= 0 && seeded && tau > 0.0f && a.runthr) ? float(NBH) / tau : 0.0f;
        }
    }
    __syncthreads();

    // v_mfma_f32_16x16x16_f16: lane l supplies A[l%16][4(l/16) .. +3] and
    // B[4(l/16) .. +3][l%16] and receives D[4(l/16) + r][l%16] in
    // accumulator r, i.e. queries 4h..4h+3 of one code: one 32-bit word of
    // the [s][code][16 queries] table row.
    const int lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 4, c16 = lane & 15;
    // wave wv builds the tables of subquantizers wv, wv + 8, ...: all 256
    // codes of one subquantizer in two halves of 8 code tiles, the B loads of
    // a half issued together
    constexpr int HT = 8;  // code tiles per half
    const _Float16 *pb = reinterpret_cast<const _Float16 *>(a.pb);
#pragma unroll 1
    for (int s = wave; s < M; s += FNT / 64) {
        f16x4 av[NMF];
#pragma unroll
        for (int f = 0; f < NMF; ++f)
            av[f] = *reinterpret_cast<const f16x4 *>(ao + (size_t(s) * FG + c16) * KP + f * 16 + 4 * h);
#pragma unroll 1
        for (int hf = 0; hf < 256 / 16 / HT; ++hf) {
            f16x4 bt[HT][NMF];
#pragma unroll
            for (int t = 0; t < HT; ++t)
#pragma unroll
                for (int f = 0; f < NMF; ++f)
                    bt[t][f] = *reinterpret_cast<const f16x4 *>(
                        pb + (size_t(s) * 256 + (hf * HT + t) * 16 + c16) * KP + f * 16 + 4 * h);
#pragma unroll
            for (int t = 0; t < HT; ++t) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int f = 0; f < NMF; ++f) acc = __builtin_amdgcn_mfma_f32_16x16x16f16(av[f], bt[t][f], acc, 0, 0, 0);
                uint32_t w = 0u;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    w = __builtin_amdgcn_cvt_pk_u8_f32(__builtin_amdgcn_fmed3f(acc[r], 0.0f, 255.0f), uint32_t(r), w);
                const int c = (hf * HT + t) * 16 + c16;
                // stored as signed bytes t - 128, the operand form of the i8 MFMA sums
                *reinterpret_cast<uint32_t *>(smem + L::t8 + (size_t(s) * 256 + c) * FG + 4 * h) = w ^ 0x80808080u;
            }
        }
    }
    __syncthreads();  // tables ready; the pool may now overwrite the A operands
}

// B operands of the table dot products, once per index:
// pb[s][c][k] = p_sc[k] / sigma_s (k < ds), 1 (k = ds), |p_sc|^2 / sigma_s^2
// (k = ds + 1), 0 after; rows of codes c >= ks are 0 (valid codes are < ks,
// so those table entries are never indexed).
template <int M>
__global__ void fast_prep_kernel(const float *__restrict__ pq, const float *__restrict__ pmax, int ks,
                                 _Float16 *__restrict__ pb) {
    constexpr int DS = D / M;
    constexpr int KP = FK<M>::KP;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // s * 256 + c
    if (i >= M * 256) return;
    const int s = i / 256, c = i % 256;
    const float sg = sigma_of(pmax[s]);
    _Float16 *row = pb + size_t(i) * KP;
    float pn = 0.0f;
    if (c < ks) {
        const float *p = pq + (int64_t(s) * ks + c) * DS;
        for (int t = 0; t < DS; ++t) {
            pn = __builtin_fmaf(p[t], p[t], pn);
            row[t] = _Float16(p[t] / sg);
        }
        row[DS] = _Float16(1.0f);
        row[DS + 1] = _Float16(pn / (sg * sg));
    } else {
        for (int t = 0; t < DS + 2; ++t) row[t] = _Float16(0.0f);
    }
    for (int t = DS + 2; t < KP; ++t) row[t] = _Float16(0.0f);
}

// Persistent: workgroups take items from a global counter.  While item i is
// scanned, the next item's index, its record and then its query and centroid
// values are fetched one stage per scan step, so the table phase of item
// i + 1 starts with its inputs in registers.
template <int M>
__global__ __launch_bounds__(FNT, 4) void fast_scan_kernel(FastArgs a) {  // 4 waves per SIMD: two workgroups per CU
    using L = FLay<M>;
    constexpr int RW = int(sizeof(FastItem) / 4);  // record words
    static_assert(RW <= 64, "a record is loaded by one wavefront");
    extern __shared__ __align__(16) unsigned char smem[];
    float *pmx = reinterpret_cast<float *>(smem + L::pmx);
    const float *delt = reinterpret_cast<const float *>(smem + L::delt);
    const uint32_t *stmp_sh = reinterpret_cast<const uint32_t *>(smem + L::stmp);  // per-slot limits
    int *qid = reinterpret_cast<int *>(smem + L::qid);
    constexpr int NJ = M / 4;               // subquantizers per lane per tile = MFMAs per tile
    constexpr int UT = 4;                   // tiles of 16 vectors per wave per step
    static_assert(UT * 16 * (FNT / 64) == FNT, "a step covers FNT vectors");
    int *pcnt = reinterpret_cast<int *>(smem + L::pcnt);
    int *item_sh = reinterpret_cast<int *>(smem + L::item);
    int *nidx_sh = reinterpret_cast<int *>(smem + L::nidx);
    FastItem *rec = reinterpret_cast<FastItem *>(smem + L::rec);
    FastItem *nrec = reinterpret_cast<FastItem *>(smem + L::nrec);
    uint2 *pool = reinterpret_cast<uint2 *>(smem + L::pool);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    if (uint32_t(uintptr_t((__attribute__((address_space(3))) unsigned char *)(smem))) != 0) {
        if (tid == 0) atomicOr(a.err, 2);
        return;
    }
    // one-hot B operand of the sums: byte (lane & 15) of lane's 16 bytes is 1
    i32x4 bsel;
    {
        const int r = lane & 15;
        const int w1 = 1 << (8 * (r & 3));
        bsel = i32x4{(r >> 2) == 0 ? w1 : 0, (r >> 2) == 1 ? w1 : 0, (r >> 2) == 2 ? w1 : 0, (r >> 2) == 3 ? w1 : 0};
    }
    const int total = *a.total_items;
    if (tid < 4) pcnt[tid] = 0;
    if (tid == 0) *item_sh = atomicAdd(a.next_item, 1);
    if (tid < M) pmx[tid] = a.pmax[tid];
    __syncthreads();
    if (*item_sh < total && tid < RW)
        reinterpret_cast<uint32_t *>(rec)[tid] = reinterpret_cast<const uint32_t *>(a.items + *item_sh)[tid];
    __syncthreads();
    ItemIn<M> in;
    if (*item_sh < total && tid < FG * M) in.load(a, rec->l, rec->q[tid / M], tid % M);

    for (;;) {
        const int item = *item_sh;
        if (item >= total) break;  // uniform
        const int64_t beg = rec->beg;
        const int len = rec->len;
        if (tid < FG) qid[tid] = rec->q[tid];
        if (tid < 3) pcnt[tid] = 0;
        __syncthreads();
        int hc[HSW];
        load_item_hist(a, qid, hc);
        build_item_tables<M>(a, smem, in, a.tau ? nullptr : rec->tau);
        item_bounds<M>(a, smem, hc);
        lds_barrier();  // the limits are final

        // prefetch of the next item: 0 index wanted, 1 record in registers,
        // 2 record in LDS, 3 inputs in registers (stage is uniform)
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
            } else if (stage == 2) {
                if (*nidx_sh < total && tid < FG * M) in.load(a, nrec->l, nrec->q[tid / M], tid % M);
            }
            ++stage;
        };

        // ---- scan on the int8 matrix cores.  A wave takes tiles of 16
        //      vectors; lane l = 16 kb + r gathers, for vector r of the tile,
        //      the 16 query bytes of subquantizers kb*NJ .. kb*NJ + NJ - 1
        //      (one ds_read_b128 each) and feeds them as the A operand of
        //      v_mfma_i32_16x16x64_i8 against a one-hot B (byte r of lane
        //      r's B is 1), so D[vector][query] sums the entries of four
        //      subquantizers per instruction.  Lane l then holds the sums of
        //      vectors 4 kb .. 4 kb + 3 of the tile for query slot r.
        const uint8_t *lc = a.codes + beg * M;
        const int r16 = lane & 15, kb = lane >> 4;
        const uint32_t lim = stmp_sh[r16];
        // a sum acc + 128 M lies in [0, 255 M], so sum < lim is acc < limi
        const int limi = int(min(lim, uint32_t(255 * M + 1))) - 128 * M;
        const uint32_t tb = uint32_t(L::t8) + uint32_t(kb * NJ) * 4096u;
        const int wv0 = wave * (UT * 16);  // first vector of this wave's tiles in a step
        // code loads run one step ahead (addresses clamped to the list, so
        // loads past its end stay in bounds)
        uint32_t cn[UT];
#pragma unroll
        for (int u = 0; u < UT; ++u) cn[u] = load_tcodes<M>(lc, min(wv0 + u * 16 + r16, max(len - 1, 0)), kb);
        int fill = 0;  // staged candidates before this step (uniform)
        int rot = 0;   // append counter of this step
        for (int base = 0; base < len; base += FNT) {
            if (tid == 0) pcnt[rot == 2 ? 0 : rot + 1] = 0;
            uint32_t cw[UT];
#pragma unroll
            for (int u = 0; u < UT; ++u) cw[u] = cn[u];
            if (base + FNT < len) {
#pragma unroll
                for (int u = 0; u < UT; ++u)
                    cn[u] = load_tcodes<M>(lc, min(base + FNT + wv0 + u * 16 + r16, len - 1), kb);
            }
            // gathers in batches of GB subquantizer steps (GB * UT reads in
            // flight), each followed by its sums
            constexpr int GB = NJ >= 2 ? 2 : 1;
            i32x4 acc[UT];
#pragma unroll
            for (int u = 0; u < UT; ++u) acc[u] = i32x4{0, 0, 0, 0};
#pragma unroll
            for (int j0 = 0; j0 < NJ; j0 += GB) {
                i32x4 xg[GB][UT];
#pragma unroll
                for (int j = 0; j < GB; ++j)
#pragma unroll
                    for (int u = 0; u < UT; ++u) {
                        const uint32_t code = (cw[u] >> (8 * (j0 + j))) & 255u;
                        xg[j][u] = *reinterpret_cast<lds_i32x4 *>(uintptr_t(tb + uint32_t(j0 + j) * 4096u + code * 16u));
                    }
#pragma unroll
                for (int j = 0; j < GB; ++j)
#pragma unroll
                    for (int u = 0; u < UT; ++u)
                        acc[u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(xg[j][u], bsel, acc[u], 0, 0, 0);
            }
            // pass bits: bit 4u + i = vector base + wv0 + 16u + 4kb + i; the
            // row test is needed only in a list's last, partial step
            uint32_t pm = 0u;
            const int vrow = base + wv0 + 4 * kb;
            if (base + FNT <= len) {  // uniform
#pragma unroll
                for (int u = 0; u < UT; ++u)
#pragma unroll
                    for (int i = 0; i < 4; ++i) pm |= uint32_t(acc[u][i] < limi) << (4 * u + i);
            } else {
                const int rem = len - vrow;
#pragma unroll
                for (int u = 0; u < UT; ++u)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        pm |= uint32_t(16 * u + i < rem && acc[u][i] < limi) << (4 * u + i);
            }
            if (__ballot(pm != 0u) != 0) {  // wave-uniform
                // wave prefix sum of the per-lane pass counts (<= 16: five bit
                // planes; two when no lane has more than three, the usual case)
                const uint32_t n = uint32_t(__popc(pm));
                uint32_t pre = 0u, tot = 0u;
                auto plane = [&](int b) {
                    const uint64_t bm = __ballot((n >> b) & 1u);
                    pre += uint32_t(__builtin_amdgcn_mbcnt_hi(uint32_t(bm >> 32),
                                                              __builtin_amdgcn_mbcnt_lo(uint32_t(bm), 0u))) << b;
                    tot += uint32_t(__popcll(bm)) << b;
                };
                plane(0);
                plane(1);
                if (__ballot(n > 3u) != 0) {  // wave-uniform
                    plane(2);
                    plane(3);
                    plane(4);
                }
                int slot = 0;
                if (lane == 0) slot = atomicAdd(pcnt + rot, int(tot));
                slot = fill + __shfl(slot, 0) + int(pre);
                const int g = r16;
                while (pm) {
                    const int bit = __ffs(pm) - 1;
                    pm &= pm - 1u;
                    // select without dynamic register indexing (bit varies per
                    // lane): a tree over the bits of bit = 4u + i
                    int cu[UT];
#pragma unroll
                    for (int u = 0; u < UT; ++u) {
                        const int a01 = (bit & 1) ? acc[u][1] : acc[u][0];
                        const int a23 = (bit & 1) ? acc[u][3] : acc[u][2];
                        cu[u] = (bit & 2) ? a23 : a01;
                    }
                    static_assert(UT == 4, "the tree below selects one of four tiles");
                    const int s01 = (bit & 4) ? cu[1] : cu[0];
                    const int s23 = (bit & 4) ? cu[3] : cu[2];
                    const int sv = (bit & 8) ? s23 : s01;
                    const uint32_t sum = uint32_t(sv + 128 * M);
                    const int v = vrow + 16 * (bit >> 2) + (bit & 3);
                    if (slot < POOL) {
                        pool[slot] = make_uint2(uint32_t(v), uint32_t(g) | (sum << 4));
                    } else {  // pool full: straight to the global buffer
                        const int q = qid[g];
                        const int pos = atomicAdd(a.cand_cnt + q, 1);
                        if (pos < a.capq)
                            a.cand[int64_t(q) * a.capq + pos] = ckey(delt[g] * float(sum), int32_t(beg + v));
                    }
                    ++slot;
                }
            }
            // Three append counters rotate over the steps: counter rot takes
            // this step's appends, counter rot + 1 (cleared at this step's
            // start) the next step's, and counter rot + 2, read by every
            // thread before this step's barrier, is free.  One barrier per
            // step then orders the appends before the reads.
            lds_barrier();  // this step's appends are done
            fill += pcnt[rot];
            rot = rot == 2 ? 0 : rot + 1;
            if (fill > POOL - FNT) {  // uniform
                flush_pool<M>(a, smem, beg, fill);
                fill = 0;
            }
            if (stage < 3) advance();
        }
        if (fill > 0) flush_pool<M>(a, smem, beg, fill);
        while (stage < 3) {
            // the previous stage (possibly the last step's, which no barrier
            // follows when nothing is left to flush) wrote nrec / nidx, which
            // this stage reads
            __syncthreads();
            advance();
        }
        __syncthreads();  // rec, qid and the pool are free; nrec is complete
        if (tid < RW) reinterpret_cast<uint32_t *>(rec)[tid] = reinterpret_cast<const uint32_t *>(nrec)[tid];
        if (tid == 0) *item_sh = *nidx_sh;
        __syncthreads();
    }
}

// One work item's tables copied out (test hook): queries 0..n-1, list l.
template <int M>
__global__ __launch_bounds__(FNT) void fast_tables_kernel(FastArgs a, int l, int n, uint8_t *t8, float *delta) {
    using L = FLay<M>;
    extern __shared__ __align__(16) unsigned char smem[];
    float *pmx = reinterpret_cast<float *>(smem + L::pmx);
    int *qid = reinterpret_cast<int *>(smem + L::qid);
    FastItem *rec = reinterpret_cast<FastItem *>(smem + L::rec);
    const int tid = threadIdx.x;
    if (tid < FG) {
        qid[tid] = tid < n ? tid : -1;
        rec->q[tid] = tid < n ? tid : -1;
        rec->tau[tid] = __builtin_huge_valf();
    }
    if (tid < M) pmx[tid] = a.pmax[tid];
    __syncthreads();
    ItemIn<M> in;
    if (tid < FG * M) in.load(a, l, rec->q[tid / M], tid % M);
    build_item_tables<M>(a, smem, in, rec->tau);
    const uint4 *src = reinterpret_cast<const uint4 *>(smem + L::t8);
    for (int i = tid; i < M * 256; i += FNT) {
        uint4 x = src[i];
        x.x ^= 0x80808080u; x.y ^= 0x80808080u; x.z ^= 0x80808080u; x.w ^= 0x80808080u;
        reinterpret_cast<uint4 *>(t8)[i] = x;
    }
    if (tid < FG) delta[tid] = reinterpret_cast<const float *>(smem + L::delt)[tid];
}

// ---- seed lists per query
__global__ void fill_i32_kernel(int *__restrict__ p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// up to six fills in one launch (grid-stride over their concatenation)
__global__ void fill_multi_kernel(FillSet f) {
    int64_t total = 0;
    for (int j = 0; j < f.count; ++j) total += f.n[j];
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
        int64_t r = i;
        int j = 0;
        while (r >= f.n[j]) r -= f.n[j++];
        f.p[j][r] = f.v[j];
    }
}

// ---- list-major grouping of (query, probe) pairs with probe rank in [rlo, rhi).
// Each workgroup aggregates its PCH pairs in an LDS histogram over the lists
// (nlist <= LBINS) and issues one global atomic per list it touches, so
// popular lists do not serialise thousands of atomics.  With more lists the
// pairs go straight to global atomics (pair_*_global_kernel); the order of a
// list's queries then follows the atomics, which changes how queries share
// work items but no result (every distance, bound and selection key is per
// query).
constexpr int LBINS = 8192;
constexpr int PCH = 2048;

__device__ inline bool pair_ok(const int32_t *probes, int64_t e, int w, int rlo, int rhi, int nlist, int *l) {
    const int r = int(e % w);
    *l = probes[e];
    return r >= rlo && r < rhi && *l >= 0 && *l < nlist;
}

__global__ __launch_bounds__(256) void pair_count_kernel(const int32_t *__restrict__ probes, int nq, int w,
                                                         int rlo, int rhi, int nlist,
                                                         int *__restrict__ cnt) {
    __shared__ int hc[LBINS];
    const int tid = threadIdx.x;
    const int64_t np = int64_t(nq) * w;
    const int64_t e0 = int64_t(blockIdx.x) * PCH;
    for (int i = tid; i < nlist; i += 256) hc[i] = 0;
    __syncthreads();
    for (int i = tid; i < PCH; i += 256) {
        const int64_t e = e0 + i;
        int l;
        if (e < np && pair_ok(probes, e, w, rlo, rhi, nlist, &l)) atomicAdd(&hc[l], 1);
    }
    __syncthreads();
    for (int i = tid; i < nlist; i += 256)
        if (hc[i]) atomicAdd(cnt + i, hc[i]);
}

__global__ __launch_bounds__(256) void pair_count_global_kernel(const int32_t *__restrict__ probes, int nq, int w,
                                                                int rlo, int rhi, int nlist,
                                                                int *__restrict__ cnt) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    int l;
    if (e < int64_t(nq) * w && pair_ok(probes, e, w, rlo, rhi, nlist, &l)) atomicAdd(cnt + l, 1);
}

__global__ __launch_bounds__(256) void pair_scatter_global_kernel(const int32_t *__restrict__ probes, int nq, int w,
                                                                  int rlo, int rhi, int nlist,
                                                                  const int *__restrict__ pair_off,
                                                                  int *__restrict__ fill, int32_t *__restrict__ pair_q) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    int l;
    if (e < int64_t(nq) * w && pair_ok(probes, e, w, rlo, rhi, nlist, &l))
        pair_q[pair_off[l] + atomicAdd(fill + l, 1)] = int32_t(e / w);
}

// exclusive scans (one workgroup) of the pair counts in list order
// (pair_off[l]) and of the work items per list in the order lorder gives
// (item_off[p] for the list lorder[p]; list order if lorder is null); items
// are numbered from *item_base (0 if null), and *total_items becomes the end
// of the numbering (item_base may equal total_items)
__global__ void pair_scan_kernel(const int *__restrict__ cnt, int nlist, int G, int *__restrict__ pair_off,
                                 int *__restrict__ item_off, int *total_items, int *__restrict__ fill,
                                 const int *item_base, const int32_t *__restrict__ lorder) {
    __shared__ int wsum_p[16], wsum_i[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int base = item_base ? *item_base : 0;  // read by every thread before the first barrier
    const int per = (nlist + 1023) / 1024;
    int sp = 0, si = 0;
    for (int j = tid * per; j < min(nlist, (tid + 1) * per); ++j) {
        sp += cnt[j];
        si += (cnt[lorder ? lorder[j] : j] + G - 1) / G;
    }
    // exclusive block scan of (sp, si): inclusive wave scans, then the wave totals
    int ip = sp, ii = si;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int op = __shfl_up(ip, off, 64), oi = __shfl_up(ii, off, 64);
        if (lane >= off) {
            ip += op;
            ii += oi;
        }
    }
    if (lane == 63) {
        wsum_p[wave] = ip;
        wsum_i[wave] = ii;
    }
    __syncthreads();
    int ap = ip - sp, ai = base + ii - si;
    for (int v = 0; v < wave; ++v) {
        ap += wsum_p[v];
        ai += wsum_i[v];
    }
    if (tid == 1023) {
        pair_off[nlist] = ap + sp;
        item_off[nlist] = ai + si;
        *total_items = ai + si;
    }
    for (int j = tid * per; j < min(nlist, (tid + 1) * per); ++j) {
        pair_off[j] = ap;
        item_off[j] = ai;
        fill[j] = 0;
        ap += cnt[j];
        ai += (cnt[lorder ? lorder[j] : j] + G - 1) / G;
    }
}

__global__ __launch_bounds__(256) void pair_scatter_kernel(const int32_t *__restrict__ probes, int nq, int w,
                                                           int rlo, int rhi, int nlist,
                                                           const int *__restrict__ pair_off, int *__restrict__ fill,
                                                           int32_t *__restrict__ pair_q) {
    __shared__ int hc[LBINS];
    __shared__ int rk[PCH];
    const int tid = threadIdx.x;
    const int64_t np = int64_t(nq) * w;
    const int64_t e0 = int64_t(blockIdx.x) * PCH;
    for (int i = tid; i < nlist; i += 256) hc[i] = 0;
    __syncthreads();
    for (int i = tid; i < PCH; i += 256) {
        const int64_t e = e0 + i;
        int l;
        rk[i] = (e < np && pair_ok(probes, e, w, rlo, rhi, nlist, &l)) ? atomicAdd(&hc[l], 1) : -1;
    }
    __syncthreads();
    for (int i = tid; i < nlist; i += 256)
        if (hc[i]) hc[i] = pair_off[i] + atomicAdd(fill + i, hc[i]);  // this workgroup's range start
    __syncthreads();
    for (int i = tid; i < PCH; i += 256) {
        if (rk[i] < 0) continue;
        const int64_t e = e0 + i;
        pair_q[hc[probes[e]] + rk[i]] = int32_t(e / w);
    }
}

// One thread per (work item, query slot); items are numbered list by list
// in lorder's order (item_off by position) from item_off[0] (the range's
// first item), so a binary search over item_off finds the item's list.
__global__ void item_expand_kernel(const int *__restrict__ pair_off, const int *__restrict__ item_off,
                                   const int64_t *__restrict__ list_off, const int32_t *__restrict__ pair_q,
                                   const float *__restrict__ tau, int nlist, int G,
                                   const int *__restrict__ total_items, int64_t max_slots,
                                   const int32_t *__restrict__ lorder, FastItem *__restrict__ items) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= max_slots) return;
    const int it = int(e / FG) + item_off[0], slot = int(e % FG);  // FG threads per record, slots >= G empty
    if (it >= *total_items) return;
    int lo = 0, hi = nlist;  // last position with item_off[p] <= it (lists without items are skipped)
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (item_off[mid] <= it) lo = mid;
        else hi = mid;
    }
    const int l = lorder ? lorder[lo] : lo;
    const int j = it - item_off[lo];
    const int pb = pair_off[l] + j * G;
    const int gc = min(G, pair_off[l + 1] - pb);
    FastItem &x = items[it];
    const int q = slot < gc ? pair_q[pb + slot] : -1;
    x.q[slot] = q;
    x.tau[slot] = q >= 0 ? (tau ? tau[q] : __builtin_huge_valf()) : -1.0f;
    if (slot == 0) {
        x.l = l;
        x.gc = gc;
        x.len = int(list_off[l + 1] - list_off[l]);
        x.pad = 0;
        x.beg = list_off[l];
        x.pad2 = 0;
    }
}

// ---- seed threshold: an upper bound of the kl-th smallest ADC distance over
// each query's seed lists.  Tables in FP32 from r = fl(q - c_l):
// T = |r_s|^2 + |p|^2 - 2 r_s.p; the kl-th value is located by a three-pass
// radix histogram (8-bit digits) of the order keys, and tau is the largest
// float of the final 24-bit bucket plus a margin 2^-14 * sum_s (|r_s| +
// max|p_s|)^2 that covers the FP32 rounding differences between these sums
// and the reference table arithmetic (each at most ~40 eps of that sum).
template <int M>
__device__ inline float seed_adc(const float *tab, const uint8_t *lc, int v) {
    uint32_t cw[M / 4];
    load_codes<M>(lc, v, cw);
    float acc = 0.0f;
#pragma unroll
    for (int s = 0; s < M; ++s) acc += tab[s * 256 + ((cw[s >> 2] >> (8 * (s & 3))) & 255u)];
    return acc;
}

// wave 0: the bin holding the need-th key (1-based) of hist[256]
__device__ inline void find_bin(const uint32_t *hist, int need, int *digit, int *below) {
    const int tid = threadIdx.x;
    if (tid < 64) {
        const uint32_t h0 = hist[4 * tid], h1 = hist[4 * tid + 1], h2 = hist[4 * tid + 2], h3 = hist[4 * tid + 3];
        const uint32_t tot = h0 + h1 + h2 + h3;
        uint32_t incl = tot;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, 64);
            if (tid >= off) incl += o;
        }
        const uint32_t excl = incl - tot;
        if (excl < uint32_t(need) && incl >= uint32_t(need)) {
            uint32_t c = excl;
            int dg;
            if (c + h0 >= uint32_t(need)) dg = 0;
            else if (c + h0 + h1 >= uint32_t(need)) { c += h0; dg = 1; }
            else if (c + h0 + h1 + h2 >= uint32_t(need)) { c += h0 + h1; dg = 2; }
            else { c += h0 + h1 + h2; dg = 3; }
            *digit = 4 * tid + dg;
            *below = int(c);
        }
    }
}

// FP32 table of query q against list l into tab (256 threads; uniform call,
// ends with a barrier); also returns sum_s (|r_s| + max|p_s|)^2 to thread 0.
template <int M>
__device__ float seed_table(const float *__restrict__ queries, const float *__restrict__ coarse,
                            const float *__restrict__ pq, const float *__restrict__ pnorm,
                            const float *__restrict__ pmax, int64_t q, int l, int ks, float *tab, float *r,
                            float *As) {
    constexpr int DS = D / M;
    const int tid = threadIdx.x;
    __syncthreads();  // the previous table is no longer read
    if (tid < D) r[tid] = queries[q * D + tid] - coarse[int64_t(l) * D + tid];
    __syncthreads();
    if (tid < M) {
        float a = 0.0f;
#pragma unroll
        for (int t = 0; t < DS; ++t) a = __builtin_fmaf(r[tid * DS + t], r[tid * DS + t], a);
        As[tid] = a;
    }
    __syncthreads();
    float mg = 0.0f;
    if (tid == 0)
        for (int s2 = 0; s2 < M; ++s2) {
            const float b = __builtin_sqrtf(As[s2]) + pmax[s2];
            mg += b * b;
        }
    if (tid < ks) {
        // four subquantizers per batch: their loads are issued together
#pragma unroll 1
        for (int s0 = 0; s0 < M; s0 += 4) {
            float p[4][DS], pn[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float4 *p4 = reinterpret_cast<const float4 *>(pq + (int64_t(s0 + u) * ks + tid) * DS);
#pragma unroll
                for (int t = 0; t < DS / 4; ++t) {
                    const float4 x = p4[t];
                    p[u][4 * t] = x.x; p[u][4 * t + 1] = x.y; p[u][4 * t + 2] = x.z; p[u][4 * t + 3] = x.w;
                }
                pn[u] = pnorm[(s0 + u) * ks + tid];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int s2 = s0 + u;
                float dot = 0.0f;
#pragma unroll
                for (int t = 0; t < DS; ++t) dot = __builtin_fmaf(r[s2 * DS + t], p[u][t], dot);
                tab[s2 * 256 + tid] = As[s2] + pn[u] - 2.0f * dot;
            }
        }
    }
    __syncthreads();
    return mg;
}

constexpr int SCAP = 8192;  // seed keys held in LDS (at most)

// SC: keys held in LDS.  A capped seed runs with the smallest SC that holds
// the cap, so that more blocks share a CU; with SC <= 2048 the launch bound
// (5 waves per SIMD, at most 102 VGPRs) lets five of them run at once and
// leaves LDS for the (query, list) grouping kernels that run beside the seed
// on a second stream (larger key buffers hold the CU to four blocks or fewer
// through their LDS).
template <int M, int SC>
__global__ __launch_bounds__(256, (SC <= 2048 ? 5 : SC <= 4096 ? 4 : 3)) void seed_tau_kernel(const float *__restrict__ queries,
                                                       const float *__restrict__ coarse,
                                                       const float *__restrict__ pq, const float *__restrict__ pnorm,
                                                       const float *__restrict__ pmax,
                                                       const int64_t *__restrict__ list_off,
                                                       const uint8_t *__restrict__ codes,
                                                       const int32_t *__restrict__ probes, int w,
                                                       int min_seed, int nlist, int ks, int kl0,
                                                       const int32_t *__restrict__ klq, int cap,
                                                       float *__restrict__ tau) {
    __shared__ float tab[M * 256];
    __shared__ float r[D];
    __shared__ float As[M];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t skeys[SC];
    __shared__ int sdig, sbelow;
    const int64_t q = blockIdx.x;
    const int tid = threadIdx.x;
    const int kl = klq ? klq[q] : kl0;
    // seed lists: the smallest ns >= min_seed whose first ns probed lists
    // hold >= kl vectors (or ns = w)
    int ns = 0;
    {
        int64_t have = 0;
        for (; ns < w; ++ns) {
            if (ns >= min_seed && have >= kl) break;  // uniform
            const int l = probes[q * w + ns];
            if (l >= 0 && l < nlist) have += list_off[l + 1] - list_off[l];
        }
    }
    // with a cap, only the first cap vectors (list prefixes in probe order, a
    // sample of the lists) enter: the kl-th smallest of a subset of the
    // probed vectors still bounds the kl-th smallest of all of them
    const int64_t lim = cap > 0 ? int64_t(cap) : (int64_t(1) << 62);
    int64_t tot = 0;
    for (int rr = 0; rr < ns; ++rr) {
        const int l = probes[q * w + rr];
        if (l >= 0 && l < nlist) tot += list_off[l + 1] - list_off[l];
    }
    tot = min(tot, lim);
    if (tot < kl) {  // uniform
        if (tid == 0) tau[q] = __builtin_huge_valf();
        return;
    }
    const bool inl = tot <= SC;  // keys computed once, kept in LDS
    float marg = 0.0f;
    if (inl) {
        int off = 0;
        for (int rr = 0; rr < ns; ++rr) {
            const int l = probes[q * w + rr];
            if (l < 0 || l >= nlist) continue;  // uniform
            if (off >= tot) break;  // uniform: the cap is reached
            const int64_t beg = list_off[l];
            const int len = int(min(list_off[l + 1] - beg, tot - off));
            const uint8_t *lc = codes + beg * M;
            // four vectors per thread per round, their code loads issued
            // together; the first round's go out before the table is built
            uint32_t cw[4][M / 4];
            if (len > 0) {
#pragma unroll
                for (int u = 0; u < 4; ++u) load_codes<M>(lc, min(tid + 256 * u, len - 1), cw[u]);
            }
            marg = fmaxf(marg, seed_table<M>(queries, coarse, pq, pnorm, pmax, q, l, ks, tab, r, As));
            for (int v0 = tid; v0 < len; v0 += 1024) {
                if (v0 != tid) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) load_codes<M>(lc, min(v0 + 256 * u, len - 1), cw[u]);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    float acc = 0.0f;
#pragma unroll
                    for (int s2 = 0; s2 < M; ++s2) acc += tab[s2 * 256 + ((cw[u][s2 >> 2] >> (8 * (s2 & 3))) & 255u)];
                    if (v0 + 256 * u < len) skeys[off + v0 + 256 * u] = fkey(acc);
                }
            }
            off += len;
        }
    }
    uint32_t prefix = 0;
    int need = kl;
    int sh0 = 24, npass = 3;
    if (inl) {
        // the radix passes start below the bits that all keys share
        __shared__ uint32_t kmn_w[4], kmx_w[4];
        __syncthreads();
        uint32_t kmn = 0xFFFFFFFFu, kmx = 0u;
        for (int i = tid; i < int(tot); i += 256) {
            kmn = min(kmn, skeys[i]);
            kmx = max(kmx, skeys[i]);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            kmn = min(kmn, uint32_t(__shfl_xor(int(kmn), off)));
            kmx = max(kmx, uint32_t(__shfl_xor(int(kmx), off)));
        }
        if ((tid & 63) == 0) {
            kmn_w[tid >> 6] = kmn;
            kmx_w[tid >> 6] = kmx;
        }
        __syncthreads();
        kmn = min(min(kmn_w[0], kmn_w[1]), min(kmn_w[2], kmn_w[3]));
        kmx = max(max(kmx_w[0], kmx_w[1]), max(kmx_w[2], kmx_w[3]));
        const int cp = kmn == kmx ? 32 : __clz(int(kmn ^ kmx));  // shared leading bits
        prefix = cp == 32 ? kmn : (cp == 0 ? 0u : (kmn >> (32 - cp)) << (32 - cp));
        sh0 = max(0, 24 - cp);  // digit = bits [sh0, sh0 + 8) sit right below the shared ones
        npass = cp == 32 ? 0 : min(3, sh0 / 8 + 1);
    }
    int sh = sh0;
    for (int pass = 0; pass < npass; ++pass, sh -= 8) {
        __syncthreads();  // keys written / previous pass's bin read
        hist[tid] = 0u;   // 256 threads, 256 bins
        __syncthreads();
        if (inl) {
            for (int i = tid; i < int(tot); i += 256) {
                const uint32_t key = skeys[i];
                if (sh + 8 >= 32 || (key >> (sh + 8)) == (prefix >> (sh + 8))) atomicAdd(&hist[(key >> sh) & 255u], 1u);
            }
        } else {
            for (int rr = 0; rr < ns; ++rr) {
                const int l = probes[q * w + rr];
                if (l < 0 || l >= nlist) continue;  // uniform
                const float mg = seed_table<M>(queries, coarse, pq, pnorm, pmax, q, l, ks, tab, r, As);
                if (pass == 0) marg = fmaxf(marg, mg);
                const int64_t beg = list_off[l];
                const int len = int(list_off[l + 1] - beg);
                const uint8_t *lc = codes + beg * M;
                for (int v = tid; v < len; v += 256) {
                    const uint32_t key = fkey(seed_adc<M>(tab, lc, v));
                    if (sh + 8 >= 32 || (key >> (sh + 8)) == (prefix >> (sh + 8))) atomicAdd(&hist[(key >> sh) & 255u], 1u);
                }
            }
        }
        __syncthreads();
        find_bin(hist, need, &sdig, &sbelow);
        __syncthreads();
        prefix = (prefix & ~(0xFFu << sh)) | (uint32_t(sdig) << sh);
        need -= sbelow;
    }
    if (tid == 0) {
        // the largest key of the final bucket (all bits below the last digit set)
        const uint32_t low = npass == 0 ? 0u : ((sh + 8) >= 32 ? 0xFFFFFFFFu : ((1u << (sh + 8)) - 1u));
        const float top = key_dist(uint64_t(prefix | low) << 32);
        tau[q] = top + marg * (1.0f / 16384.0f);
    }
}

// ---- per query: the kl best (approximate distance, list position) keys of
// the candidates (in any order; the rerank sorts).  A 256-bin histogram of
// the keys' distance bits over [min, max] finds the bin b holding the kl-th
// key; keys of lower bins are taken whole, and the few keys of bin b go to
// LDS, where their ranks (up to 256 keys) or an exact radix select pick the
// missing ones.  When bin b holds more than SELB keys (heavy ties), a radix
// select over all keys is used.
constexpr int SELB = 1024;

struct ShortShared {
    SelectShared sh;
    uint64_t buf[SELB];
    uint32_t hist[256];
    uint32_t wmn[4], wmx[4];
    int sbin, sbelow, cnt, nb, eqt;
};

// The kl best keys of query q's candidates into out[0..kl) (list positions,
// -1 past the candidates; any order), by the whole workgroup.
__device__ void select_shortlist(int kl, const uint64_t *__restrict__ cand, const int *__restrict__ cand_cnt,
                                 int capq, int q, int32_t *out, ShortShared &ss) {
    SelectShared &sh = ss.sh;
    uint64_t *buf = ss.buf;
    uint32_t *hist = ss.hist;
    uint32_t *wmn = ss.wmn, *wmx = ss.wmx;
    int &sbin = ss.sbin, &sbelow = ss.sbelow, &cnt = ss.cnt, &nb = ss.nb, &eqt = ss.eqt;
    const int tid = threadIdx.x;
    const uint64_t *C = cand + int64_t(q) * capq;
    // up to 2048 candidates are read once, into registers (more are re-read);
    // the first 1024 loads go out together with the count's (entries past
    // it are dropped), the rest once the count says they exist
    constexpr int FR = 8, FS = 4;
    uint64_t rk[FR];
#pragma unroll
    for (int e = 0; e < FS; ++e) {
        const int i = tid + 256 * e;
        rk[e] = i < capq ? C[i] : kNoKey;
    }
    const int n = min(cand_cnt[q], capq);
    if (n <= kl) {  // uniform: every candidate is in
        for (int i = tid; i < kl; i += 256) out[i] = i < n ? key_id(C[i]) : -1;
        return;
    }
    const bool inreg = n <= 256 * FR;  // uniform
#pragma unroll
    for (int e = 0; e < FR; ++e) {
        const int i = tid + 256 * e;
        if (!inreg || i >= n) rk[e] = kNoKey;
        else if (e >= FS) rk[e] = C[i];
    }
    uint32_t mn = 0xFFFFFFFFu, mx = 0u;
    if (inreg) {
#pragma unroll
        for (int e = 0; e < FR; ++e) {
            if (rk[e] == kNoKey) continue;
            const uint32_t h = uint32_t(rk[e] >> 32);
            mn = min(mn, h);
            mx = max(mx, h);
        }
    } else {
        for (int i = tid; i < n; i += 256) {
            const uint32_t h = uint32_t(C[i] >> 32);
            mn = min(mn, h);
            mx = max(mx, h);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, uint32_t(__shfl_xor(int(mn), off)));
        mx = max(mx, uint32_t(__shfl_xor(int(mx), off)));
    }
    if ((tid & 63) == 0) {
        wmn[tid >> 6] = mn;
        wmx[tid >> 6] = mx;
    }
    hist[tid] = 0u;
    if (tid == 0) {
        cnt = 0;
        nb = 0;
        eqt = 0;
    }
    __syncthreads();
    mn = min(min(wmn[0], wmn[1]), min(wmn[2], wmn[3]));
    mx = max(max(wmx[0], wmx[1]), max(wmx[2], wmx[3]));
    // any monotone map of the distance bits to bins keeps the order
    const float scale = 256.0f / (float(mx - mn) + 1.0f);
    auto bin = [&](uint64_t k) { return min(255, int(float(uint32_t(k >> 32) - mn) * scale)); };
    if (inreg) {
#pragma unroll
        for (int e = 0; e < FR; ++e)
            if (rk[e] != kNoKey) atomicAdd(&hist[bin(rk[e])], 1u);
    } else {
        for (int i = tid; i < n; i += 256) atomicAdd(&hist[bin(C[i])], 1u);
    }
    __syncthreads();
    if (tid < 64) {
        const uint32_t h0 = hist[4 * tid], h1 = hist[4 * tid + 1], h2 = hist[4 * tid + 2], h3 = hist[4 * tid + 3];
        const uint32_t tot = h0 + h1 + h2 + h3;
        uint32_t incl = tot;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, 64);
            if (tid >= off) incl += o;
        }
        const uint32_t excl = incl - tot;
        const uint32_t need = uint32_t(kl);
        if (excl < need && incl >= need) {
            uint32_t c = excl;
            int dg;
            if (c + h0 >= need) dg = 0;
            else if (c + h0 + h1 >= need) { c += h0; dg = 1; }
            else if (c + h0 + h1 + h2 >= need) { c += h0 + h1; dg = 2; }
            else { c += h0 + h1 + h2; dg = 3; }
            sbin = 4 * tid + dg;
            sbelow = int(c);
        }
    }
    __syncthreads();
    const int b = sbin, below = sbelow;
    if (int(hist[b]) > SELB) {  // uniform: heavy ties, radix select over all keys
        auto keyfn = [&](int i) -> uint64_t { return C[i]; };
        const SelectResult r = block_select<256>(keyfn, n, kl, sh);
        for (int i = tid; i < n; i += 256) {
            const uint64_t key = C[i];
            bool take = key < r.thr;
            if (!take && r.eq > 0 && key == r.thr) take = atomicAdd(&eqt, 1) < r.eq;
            if (take) {
                const int pos = atomicAdd(&cnt, 1);
                if (pos < kl) out[pos] = key_id(key);
            }
        }
        return;
    }
    auto place = [&](uint64_t key) {
        const int bk = bin(key);
        if (bk < b) out[atomicAdd(&cnt, 1)] = key_id(key);
        else if (bk == b) buf[atomicAdd(&nb, 1)] = key;
    };
    if (inreg) {
#pragma unroll
        for (int e = 0; e < FR; ++e)
            if (rk[e] != kNoKey) place(rk[e]);
    } else {
        for (int i = tid; i < n; i += 256) place(C[i]);
    }
    __syncthreads();
    const int m = nb, need = kl - below;
    if (m <= 256) {  // uniform: the few keys of bin b by rank, one pass
        if (tid < m) {
            const uint64_t key = buf[tid];
            int r = 0;
            for (int i = 0; i < m; ++i) {
                const uint64_t o = buf[i];
                r += o < key || (o == key && i < tid);
            }
            if (r < need) out[atomicAdd(&cnt, 1)] = key_id(key);  // cnt starts at `below`
        }
        return;
    }
    auto bkey = [&](int i) -> uint64_t { return buf[i]; };
    const SelectResult r = block_select<256>(bkey, m, need, sh);
    for (int i = tid; i < m; i += 256) {
        const uint64_t key = buf[i];
        bool take = key < r.thr;
        if (!take && r.eq > 0 && key == r.thr) take = atomicAdd(&eqt, 1) < r.eq;
        if (take) out[atomicAdd(&cnt, 1)] = key_id(key);  // cnt starts at `below`
    }
}

__global__ __launch_bounds__(256) void fast_select_kernel(int kl, const uint64_t *__restrict__ cand,
                                                          const int *__restrict__ cand_cnt, int capq,
                                                          int32_t *__restrict__ short_ids) {
    __shared__ ShortShared ss;
    const int q = blockIdx.x;
    select_shortlist(kl, cand, cand_cnt, capq, q, short_ids + int64_t(q) * kl, ss);
}

// Shortlist and exact rerank in one workgroup per query (kl <= 256): the
// list positions stay in LDS instead of a round trip through HBM.
template <int MODE, typename T>
__global__ __launch_bounds__(256) void fast_select_rerank_kernel(int kl, const uint64_t *__restrict__ cand,
                                                                 const int *__restrict__ cand_cnt, int capq,
                                                                 const T *__restrict__ rows, int64_t nrows,
                                                                 const int32_t *__restrict__ list_ids,
                                                                 const float *__restrict__ queries, int d,
                                                                 const double *__restrict__ qnorm, int k,
                                                                 int32_t *__restrict__ out_ids,
                                                                 float *__restrict__ out_dis) {
    __shared__ ShortShared ss;
    __shared__ RerankShared rs;
    __shared__ int32_t sl[256];
    const int q = blockIdx.x;
    select_shortlist(kl, cand, cand_cnt, capq, q, sl, ss);
    __syncthreads();
    const int32_t e = int(threadIdx.x) < kl ? sl[threadIdx.x] : -1;
    rerank_block<MODE, T, 256>(rows, nrows, list_ids, queries, d, qnorm, q, e, k, out_ids, out_dis, rs);
}

inline unsigned nblk(int64_t n, int t) { return unsigned((n + t - 1) / t); }

template <int M>
hipError_t set_lds_attr(const void *fn) {
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(FLay<M>::total));
}

template <int M>
hipError_t launch_fast_m(const FastArgs &a, int grid, hipStream_t s) {
    static bool attr = false;
    if (!attr) {
        hipError_t e = set_lds_attr<M>(reinterpret_cast<const void *>(&fast_scan_kernel<M>));
        if (e != hipSuccess) return e;
        attr = true;
    }
    fast_scan_kernel<M><<<grid, FNT, FLay<M>::total, s>>>(a);
    return hipGetLastError();
}

template <int M>
hipError_t launch_tables_m(const FastArgs &a, int l, int n, uint8_t *t8, float *delta, hipStream_t s) {
    static bool attr = false;
    if (!attr) {
        hipError_t e = set_lds_attr<M>(reinterpret_cast<const void *>(&fast_tables_kernel<M>));
        if (e != hipSuccess) return e;
        attr = true;
    }
    fast_tables_kernel<M><<<1, FNT, FLay<M>::total, s>>>(a, l, n, t8, delta);
    return hipGetLastError();
}

}  // namespace

int fast_hist_bins() { return NBH; }
int fast_group_size() { return FG; }

size_t fast_pb_bytes(int m) {
    switch (m) {
        case 8: return size_t(8) * 256 * FK<8>::KP * 2;
        case 16: return size_t(16) * 256 * FK<16>::KP * 2;
        default: return 0;
    }
}

hipError_t launch_fast_prep(const float *pq, const float *pmax, int m, int ks, void *pb, hipStream_t s) {
    switch (m) {
        case 8:
            fast_prep_kernel<8><<<nblk(8 * 256, 256), 256, 0, s>>>(pq, pmax, ks, static_cast<_Float16 *>(pb));
            break;
        case 16:
            fast_prep_kernel<16><<<nblk(16 * 256, 256), 256, 0, s>>>(pq, pmax, ks, static_cast<_Float16 *>(pb));
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}


hipError_t launch_fill_i32(int *p, int n, int v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    fill_i32_kernel<<<nblk(n, 256), 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_multi(const FillSet &f, hipStream_t s) {
    if (f.count < 0 || f.count > 6) return hipErrorInvalidValue;
    int64_t total = 0;
    for (int j = 0; j < f.count; ++j) {
        if (f.n[j] < 0) return hipErrorInvalidValue;
        total += f.n[j];
    }
    if (total == 0) return hipSuccess;
    fill_multi_kernel<<<unsigned(std::min<int64_t>(nblk(total, 256), 2048)), 256, 0, s>>>(f);
    return hipGetLastError();
}

hipError_t launch_pairs(const int32_t *probes, int nq, int w, int rlo, int rhi, const int64_t *list_off,
                        const float *tau, int nlist, int G, int *cnt, int *pair_off, int *item_off, int *total_items,
                        const int *item_base, const int32_t *lorder, int *fill, int32_t *pair_q, FastItem *items,
                        hipStream_t s) {
    if (nlist < 1 || nlist > kMaxPairLists || G < 1 || G > FG) return hipErrorInvalidValue;
    rlo = max(rlo, 0);
    rhi = min(rhi, w);
    const int64_t np = int64_t(nq) * w;
    const int64_t npr = int64_t(nq) * max(rhi - rlo, 0);  // pairs of the rank range, at most
    fill_i32_kernel<<<nblk(nlist, 256), 256, 0, s>>>(cnt, nlist, 0);
    if (nlist <= LBINS)
        pair_count_kernel<<<nblk(np, PCH), 256, 0, s>>>(probes, nq, w, rlo, rhi, nlist, cnt);
    else
        pair_count_global_kernel<<<nblk(np, 256), 256, 0, s>>>(probes, nq, w, rlo, rhi, nlist, cnt);
    pair_scan_kernel<<<1, 1024, 0, s>>>(cnt, nlist, G, pair_off, item_off, total_items, fill, item_base, lorder);
    if (nlist <= LBINS)
        pair_scatter_kernel<<<nblk(np, PCH), 256, 0, s>>>(probes, nq, w, rlo, rhi, nlist, pair_off, fill, pair_q);
    else
        pair_scatter_global_kernel<<<nblk(np, 256), 256, 0, s>>>(probes, nq, w, rlo, rhi, nlist, pair_off, fill,
                                                                 pair_q);
    const int64_t max_slots = (npr / G + nlist) * FG;  // >= FG * items of the range
    item_expand_kernel<<<nblk(max_slots, 256), 256, 0, s>>>(pair_off, item_off, list_off, pair_q, tau, nlist, G,
                                                          total_items, max_slots, lorder, items);
    return hipGetLastError();
}

template <int M, int SC>
hipError_t launch_seed_m(const SeedArgs &a, hipStream_t s) {
    seed_tau_kernel<M, SC><<<a.nq, 256, 0, s>>>(a.queries, a.coarse, a.pq, a.pnorm, a.pmax, a.list_off, a.codes,
                                                 a.probes, a.w, a.min_seed, a.nlist, a.ks, a.kl, a.klq, a.cap, a.tau);
    return hipGetLastError();
}

hipError_t launch_seed_tau(const SeedArgs &a, int m, hipStream_t s) {
    if (a.nq <= 0) return hipSuccess;
    if (a.kl < 1) return hipErrorInvalidValue;
    // a cap must hold the kl-th key: cap >= kl (per-query klq[q] are <= kl)
    if (a.cap > 0 && (a.cap < a.kl || a.cap > SCAP)) return hipErrorInvalidValue;
    // the smallest key buffer that holds the cap: less LDS, more blocks per CU
    int sc = SCAP;
    for (int c : {1024, 2048, 4096})
        if (a.cap > 0 && a.cap <= c && sc == SCAP) sc = c;
    switch (m * 10000 + sc) {
        case 81024: return launch_seed_m<8, 1024>(a, s);
        case 82048: return launch_seed_m<8, 2048>(a, s);
        case 84096: return launch_seed_m<8, 4096>(a, s);
        case 88192: return launch_seed_m<8, SCAP>(a, s);
        case 161024: return launch_seed_m<16, 1024>(a, s);
        case 162048: return launch_seed_m<16, 2048>(a, s);
        case 164096: return launch_seed_m<16, 4096>(a, s);
        case 168192: return launch_seed_m<16, SCAP>(a, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_scan(const FastArgs &a, int m, int grid, hipStream_t s) {
    switch (m) {
        case 8: return launch_fast_m<8>(a, grid, s);
        case 16: return launch_fast_m<16>(a, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_tables(const FastArgs &a, int m, int l, int n, uint8_t *t8, float *delta, hipStream_t s) {
    if (n < 1 || n > FG) return hipErrorInvalidValue;
    switch (m) {
        case 8: return launch_tables_m<8>(a, l, n, t8, delta, s);
        case 16: return launch_tables_m<16>(a, l, n, t8, delta, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_select_rerank(int kl, const uint64_t *cand, const int *cand_cnt, int capq, int nq,
                                     const void *rows, bool rows_u8, bool list_ordered, int64_t nrows,
                                     const int32_t *list_ids, const float *queries, int d, const double *qnorm, int k,
                                     int32_t *out_ids, float *out_dis, hipStream_t s) {
    if (kl < 1 || kl > 256 || k < 1 || k > kl || d > 512) return hipErrorInvalidValue;
    if (nq <= 0) return hipSuccess;
    if (rows_u8 && list_ordered)
        fast_select_rerank_kernel<2, uint8_t><<<nq, 256, 0, s>>>(kl, cand, cand_cnt, capq,
                                                                 static_cast<const uint8_t *>(rows), nrows, list_ids,
                                                                 queries, d, qnorm, k, out_ids, out_dis);
    else if (list_ordered)
        fast_select_rerank_kernel<2, float><<<nq, 256, 0, s>>>(kl, cand, cand_cnt, capq,
                                                               static_cast<const float *>(rows), nrows, list_ids,
                                                               queries, d, qnorm, k, out_ids, out_dis);
    else if (!rows_u8)
        fast_select_rerank_kernel<1, float><<<nq, 256, 0, s>>>(kl, cand, cand_cnt, capq,
                                                               static_cast<const float *>(rows), nrows, list_ids,
                                                               queries, d, qnorm, k, out_ids, out_dis);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_fast_select(int kl, const uint64_t *cand, const int *cand_cnt, int capq, int nq,
                              int32_t *short_ids, hipStream_t s) {
    if (kl < 1 || kl > 1024) return hipErrorInvalidValue;
    if (nq <= 0) return hipSuccess;
    fast_select_kernel<<<nq, 256, 0, s>>>(kl, cand, cand_cnt, capq, short_ids);
    return hipGetLastError();
}

}  // namespace mivf
