// Table types of the batched reduced-precision scan built from FP32 tables
// (mivf_search_fast with option "fast_tables" = 1 or 2).  The work items are
// those of fast_scan_kernel: (list l, up to 16 queries that probe l), so the
// list's codes are read once per 16 queries.  Per item the 16 tables are first
// computed in FP32 with the reference arithmetic (the exact kernels' table
// formula: T = fl(fl(-2 dot) + fl(|p|^2_f64 + |r_s|^2))), then stored as
//
//   1 (FP16):  e = fp16(T * 2^-E_g), where E_g >= 0 is the smallest power of
//              two that keeps the largest possible entry of query g,
//              max_s (|r_s| + sigma_s)^2, below 2^15.  The approximate ADC is
//              2^E_g * sum_s e_s (FP32 sum).  Rounding to the nearest FP16 moves
//              an entry by at most 2^-11 of itself, so the approximate ADC is
//              within 2^-11 ADC (plus the FP32 sums' rounding) of the exact one.
//   2 (8-bit, per-subquantizer offset and scale): for each query and
//              subquantizer lo_s = min_c T_s(c), hi_s = max_c T_s(c),
//              t = round((T - lo_s) * 255 / (hi_s - lo_s)); the approximate ADC
//              is sum_s lo_s + sum_s t_s (hi_s - lo_s) / 255, within
//              sum_s (hi_s - lo_s) / 510 of the exact one.
//
// A vector is kept for query g when its approximate ADC is <= tau_g plus that
// bound (and a relative 2^-20 for the FP32 sums), so every vector whose exact
// ADC is <= tau_g (the seed bound of the query's kl-th distance) is kept, as
// in the 8-bit matrix-core scan, and when it is <= the query's running bound
// R_q (at least kl candidates already have an approximate distance <= R_q, so
// the shortlist of the kl smallest is the same with or without it); the same
// shortlist select and exact rerank follow.  Each thread scans one vector per step for all 16 queries: one or
// two 16-byte LDS reads per subquantizer, FP32 sums on the VALU.
#include "select_dev.h"

namespace mivf {
namespace {

// row stride of the running-bound histograms (FastArgs::rhist), shared with
// the 8-bit matrix-core scan
constexpr int kHistStride = 64;
__device__ inline int fast_hist_bins_dev() { return kHistStride; }

constexpr int D = 128;
constexpr int VG = 16;      // queries per work item
constexpr int VNT = 512;    // threads per workgroup
constexpr int VPOOL = 1024; // staged candidates
constexpr int VNBH = 32;    // bins of the running-bound histograms (the first VNBH of each row)

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const u32x4 lds_u32x4;
typedef __attribute__((address_space(3))) const f32x4 lds_f32x4;

template <int M, int TT>
struct VLay {
    static constexpr int EB = TT == 1 ? 2 : 1;                         // bytes per entry
    static constexpr size_t tab = 0;                                    // [M][256][VG] entries
    static constexpr size_t pool = tab + size_t(M) * 256 * VG * EB;     // uint2 [VPOOL] (scan phase)
    static constexpr size_t R = pool;                                   // float [VG][D] (table phase)
    static_assert(size_t(VG) * D * 4 <= size_t(VPOOL) * 8, "the residuals must fit in the pool");
    static constexpr size_t prank = pool + size_t(VPOOL) * 8;           // uint16 [VPOOL]
    static constexpr size_t rn = prank + size_t(VPOOL) * 2;             // float [VG][M]
    static constexpr size_t lhist = rn + size_t(VG) * M * 4;            // int [VG][VNBH] flush histogram
    static constexpr size_t hsc = lhist + size_t(VG) * VNBH * 4;        // float [VG] histogram scale VNBH / tau (0: none)
    static constexpr size_t isc = hsc + VG * 4;                         // float [M][VG] (hi - lo) / 255
    static constexpr size_t lov = isc + size_t(M) * VG * 4;             // float [M][VG] lo_s
    static constexpr size_t mul = lov + size_t(M) * VG * 4;             // float [VG] 2^E_g (FP16) / 1
    static constexpr size_t base = mul + VG * 4;                        // float [VG] sum_s lo_s (8-bit) / 0
    static constexpr size_t lim = base + VG * 4;                        // float [VG] pass limit (-inf: empty slot)
    static constexpr size_t qid = lim + VG * 4;                         // int [VG]
    static constexpr size_t gcnt = qid + VG * 4;                        // int [VG]
    static constexpr size_t gbase = gcnt + VG * 4;                      // int [VG]
    static constexpr size_t rec = gbase + VG * 4;                       // FastItem: the current item
    static constexpr size_t nrec = rec + sizeof(FastItem);              // FastItem: the next item
    static constexpr size_t pcnt = nrec + sizeof(FastItem);             // int [4] rotating append counters
    static constexpr size_t item = pcnt + 16;                           // int
    static constexpr size_t nidx = item + 4;                            // int
    static constexpr size_t taus = nidx + 4;                            // float [VG] seed bounds of the slots
    static constexpr size_t total = taus + VG * 4;
};

template <int M>
__device__ inline void vload_codes(const uint8_t *lc, int v, uint32_t (&cw)[M / 4]) {
    if constexpr (M == 16) {
        const uint4 c4 = *reinterpret_cast<const uint4 *>(lc + uint32_t(v) * 16u);
        cw[0] = c4.x; cw[1] = c4.y; cw[2] = c4.z; cw[3] = c4.w;
    } else {
        const uint2 c2 = *reinterpret_cast<const uint2 *>(lc + uint32_t(v) * 8u);
        cw[0] = c2.x; cw[1] = c2.y;
    }
}

__device__ inline float half_lo(uint32_t w) { return float(__builtin_bit_cast(_Float16, uint16_t(w & 0xFFFFu))); }
__device__ inline float half_hi(uint32_t w) { return float(__builtin_bit_cast(_Float16, uint16_t(w >> 16))); }

// The 16 tables of an item (queries qid[0..15], list l) and the per-query
// constants mul, base and lim.  taus: the queries' seed bounds (per slot).
// Uniform call; ends after a barrier with the tables complete.
template <int M, int TT>
__device__ void vtables(const FastArgs &a, unsigned char *smem, int l, const float *taus) {
    constexpr int DS = D / M;
    constexpr int GC = (TT == 2 && M == 16) ? 2 : 32 / DS;  // queries per register-resident chunk (32 residual values, 16 for the 8-bit form at m = 16)
    constexpr int E = M * 256 / VNT;  // codes per thread
    constexpr int CPS = 256 / E;      // threads per subquantizer (within one wavefront)
    static_assert(E * VNT == M * 256 && CPS <= 64 && 64 % CPS == 0, "one subquantizer per thread group");
    using L = VLay<M, TT>;
    float *R = reinterpret_cast<float *>(smem + L::R);
    float *rn = reinterpret_cast<float *>(smem + L::rn);
    float *isc = reinterpret_cast<float *>(smem + L::isc);
    float *lov = reinterpret_cast<float *>(smem + L::lov);
    float *mul = reinterpret_cast<float *>(smem + L::mul);
    float *basev = reinterpret_cast<float *>(smem + L::base);
    float *lim = reinterpret_cast<float *>(smem + L::lim);
    const int *qid = reinterpret_cast<const int *>(smem + L::qid);
    const int tid = threadIdx.x;
    for (int e = tid; e < VG * D; e += VNT) {
        const int g = e / D, t = e % D;
        const int q = qid[g];
        R[e] = q >= 0 ? a.queries[int64_t(q) * D + t] - a.coarse[int64_t(l) * D + t] : 0.0f;
    }
    __syncthreads();
    if (tid < VG * M) {
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
    if (TT == 1 && tid < VG) {
        // the largest possible entry of the query: max_s (|r_s| + sigma_s)^2
        float bound = 0.0f;
        for (int s = 0; s < M; ++s) {
            const float b = __builtin_sqrtf(rn[tid * M + s]) + a.pmax[s];
            bound = fmaxf(bound, b * b);
        }
        float m2 = 1.0f;
        for (int e = 0; e < 100 && bound * 1.001f >= 32768.0f * m2; ++e) m2 *= 2.0f;  // exact powers of two
        mul[tid] = m2;
    }
    __syncthreads();
    const int s = tid / CPS, c0 = tid % CPS;
    const int lane = tid & 63;
#pragma unroll 1
    for (int g0 = 0; g0 < VG; g0 += GC) {
        float r[GC][DS];
        double rn64[GC];
        float inv[GC];
#pragma unroll
        for (int g = 0; g < GC; ++g) {
#pragma unroll
            for (int i = 0; i < DS; ++i) r[g][i] = R[(g0 + g) * D + s * DS + i];
            rn64[g] = double(rn[(g0 + g) * M + s]);
            inv[g] = TT == 1 ? 1.0f / mul[g0 + g] : 1.0f;  // a power of two: exact
        }
        // the GC entries T of code c = c0 + CPS j (NaN for codes >= ks:
        // never indexed, never a minimum or maximum)
        auto entries = [&](int j, float (&t)[GC]) {
            const int c = c0 + CPS * j;
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
                    t[g] = __builtin_fmaf(-2.0f, dot, cc);
                }
            } else {
#pragma unroll
                for (int g = 0; g < GC; ++g) t[g] = __builtin_nanf("");
            }
        };
        if constexpr (TT == 1) {
#pragma unroll 1
            for (int j = 0; j < E; ++j) {
                const int c = c0 + CPS * j;
                float t[GC];
                entries(j, t);
                uint32_t w[GC / 2];
#pragma unroll
                for (int h = 0; h < GC / 2; ++h) {
                    const _Float16 x0 = c < a.ks ? _Float16(t[2 * h] * inv[2 * h]) : _Float16(0.0f);
                    const _Float16 x1 = c < a.ks ? _Float16(t[2 * h + 1] * inv[2 * h + 1]) : _Float16(0.0f);
                    w[h] = uint32_t(__builtin_bit_cast(uint16_t, x0)) | (uint32_t(__builtin_bit_cast(uint16_t, x1)) << 16);
                }
                unsigned char *dst = smem + L::tab + (size_t(s * 256 + c) * VG + g0) * 2;
                if constexpr (GC == 4) *reinterpret_cast<uint2 *>(dst) = make_uint2(w[0], w[1]);
                else *reinterpret_cast<uint32_t *>(dst) = w[0];
            }
        } else {
            // pass 1: per (query, subquantizer) the min and max over the
            // codes of the CPS threads that hold subquantizer s (one
            // wavefront or half of one); pass 2 recomputes the entries and
            // quantizes them (cheaper than keeping E * GC values live)
            float lo[GC], hi[GC];
#pragma unroll
            for (int g = 0; g < GC; ++g) {
                lo[g] = __builtin_huge_valf();
                hi[g] = -__builtin_huge_valf();
            }
#pragma unroll 1
            for (int j = 0; j < E; ++j) {
                float t[GC];
                entries(j, t);
#pragma unroll
                for (int g = 0; g < GC; ++g) {
                    lo[g] = fminf(lo[g], t[g]);  // fminf ignores the NaN of codes >= ks
                    hi[g] = fmaxf(hi[g], t[g]);
                }
            }
#pragma unroll
            for (int g = 0; g < GC; ++g) {
#pragma unroll
                for (int off = CPS / 2; off > 0; off >>= 1) {
                    lo[g] = fminf(lo[g], __shfl_xor(lo[g], off, 64));
                    hi[g] = fmaxf(hi[g], __shfl_xor(hi[g], off, 64));
                }
            }
#pragma unroll 1
            for (int j = 0; j < E; ++j) {
                const int c = c0 + CPS * j;
                float tv[GC];
                entries(j, tv);
                uint32_t wb = 0u;
#pragma unroll
                for (int g = 0; g < GC; ++g) {
                    const float rg = hi[g] - lo[g];
                    const float sc = rg > 0.0f ? 255.0f / rg : 0.0f;
                    const float t = c < a.ks ? fminf(fmaxf(__builtin_rintf((tv[g] - lo[g]) * sc), 0.0f), 255.0f)
                                             : 0.0f;
                    wb |= uint32_t(t) << (8 * g);
                }
                unsigned char *dst = smem + L::tab + size_t(s * 256 + c) * VG + g0;
                if constexpr (GC == 4) *reinterpret_cast<uint32_t *>(dst) = wb;
                else *reinterpret_cast<uint16_t *>(dst) = uint16_t(wb);
            }
            if ((lane & (CPS - 1)) == 0) {
#pragma unroll
                for (int g = 0; g < GC; ++g) {
                    const float rg = hi[g] - lo[g];
                    isc[s * VG + g0 + g] = rg > 0.0f ? rg / 255.0f : 0.0f;
                    lov[s * VG + g0 + g] = rg >= 0.0f ? lo[g] : 0.0f;
                }
            }
        }
    }
    __syncthreads();
    if (tid < VG) {
        const int g = tid;
        const float tau = taus[g];
        float bnd = 0.0f, slack = 0.0f, b0 = 0.0f;
        for (int s2 = 0; s2 < M; ++s2) {
            const float b = __builtin_sqrtf(rn[g * M + s2]) + a.pmax[s2];
            bnd += b * b;  // >= every ADC of the query's table
            if (TT == 2) {
                b0 = b0 + lov[s2 * VG + g];
                slack += 0.5002f * isc[s2 * VG + g];  // half a step, and the FP32 rounding of (T - lo) * 255 / range
            }
        }
        if (TT == 2) {
            mul[g] = 1.0f;
            basev[g] = b0;
        } else {
            basev[g] = 0.0f;
        }
        const float rel = TT == 1 ? 0x1.0p-10f : 0.0f;
        const bool seeded = tau < 3.0e38f;
        float lm;
        if (qid[g] < 0) lm = -__builtin_huge_valf();  // empty slot: nothing passes
        else if (!seeded) lm = __builtin_huge_valf();  // no seed bound: keep everything
        else lm = tau * (1.0f + rel) + slack + 0x1.0p-20f * (fabsf(tau) + bnd);
        if (qid[g] >= 0 && a.runthr) {
            // running bound R: >= kl candidates of the query already have an
            // approximate distance <= R, so a vector above R cannot enter its
            // shortlist of the kl smallest
            lm = fminf(lm, a.runthr[qid[g]]);
        }
        lim[g] = lm;
        float *hsc = reinterpret_cast<float *>(smem + L::hsc);
        hsc[g] = (qid[g] >= 0 && seeded && tau > 0.0f && a.runthr) ? float(VNBH) / tau : 0.0f;
    }
    __syncthreads();  // tables and constants ready; the pool may now overwrite R and rn
}

// Approximate ADC of vector v (codes cw) for the 16 query slots.  The
// subquantizers go in groups of SU per loop trip (the code words rotate), so
// only SU subquantizers' reads are in flight at a time.
template <int M, int TT>
__device__ inline void vadc(unsigned char *smem, const uint32_t (&cw)[M / 4], float (&acc)[VG]) {
    using L = VLay<M, TT>;
    constexpr int SU = TT == 1 ? 4 : (M == 16 ? 1 : 2);
#pragma unroll
    for (int g = 0; g < VG; ++g) acc[g] = 0.0f;
    uint32_t words[M / 4];
#pragma unroll
    for (int i = 0; i < M / 4; ++i) words[i] = cw[i];
#pragma unroll 1
    for (int s0 = 0; s0 < M; s0 += SU) {
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int s = s0 + u;
            const uint32_t code = (words[0] >> (8 * ((s0 + u) & 3))) & 255u;
            if constexpr (TT == 1) {
                const uint32_t row = uint32_t(L::tab) + (uint32_t(s) * 256u + code) * 32u;
                const u32x4 x0 = *reinterpret_cast<lds_u32x4 *>(uintptr_t(row));
                const u32x4 x1 = *reinterpret_cast<lds_u32x4 *>(uintptr_t(row + 16u));
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    acc[2 * i] = acc[2 * i] + half_lo(x0[i]);
                    acc[2 * i + 1] = acc[2 * i + 1] + half_hi(x0[i]);
                    acc[8 + 2 * i] = acc[8 + 2 * i] + half_lo(x1[i]);
                    acc[8 + 2 * i + 1] = acc[8 + 2 * i + 1] + half_hi(x1[i]);
                }
            } else {
                const uint32_t row = uint32_t(L::tab) + (uint32_t(s) * 256u + code) * 16u;
                const u32x4 x = *reinterpret_cast<lds_u32x4 *>(uintptr_t(row));
                const uint32_t sc = uint32_t(L::isc) + uint32_t(s) * VG * 4u;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const f32x4 f = *reinterpret_cast<lds_f32x4 *>(uintptr_t(sc + 16u * i));  // same address in every lane
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[4 * i + j] = __builtin_fmaf(float((x[i] >> (8 * j)) & 255u), f[j], acc[4 * i + j]);
                }
            }
        }
        if (((s0 + SU) & 3) == 0) {  // next code word
#pragma unroll
            for (int i = 0; i + 1 < M / 4; ++i) words[i] = words[i + 1];
        }
    }
}

// Staged candidates -> per-query global buffers as (approximate distance,
// list position) keys.  Entry: x = position in the list | g << 28, y = distance bits.
template <int M, int TT>
__device__ void vflush(const FastArgs &a, unsigned char *smem, int64_t beg, int filled) {
    using L = VLay<M, TT>;
    const uint2 *pool = reinterpret_cast<const uint2 *>(smem + L::pool);
    uint16_t *prank = reinterpret_cast<uint16_t *>(smem + L::prank);
    int *gcnt = reinterpret_cast<int *>(smem + L::gcnt);
    int *gbase = reinterpret_cast<int *>(smem + L::gbase);
    int *lhist = reinterpret_cast<int *>(smem + L::lhist);
    const float *hsc = reinterpret_cast<const float *>(smem + L::hsc);
    const int *qid = reinterpret_cast<const int *>(smem + L::qid);
    const int tid = threadIdx.x;
    const int n = min(filled, VPOOL);
    if (tid < VG) gcnt[tid] = 0;
    for (int i = tid; i < VG * VNBH; i += VNT) lhist[i] = 0;
    lds_barrier();
    for (int i = tid; i < n; i += VNT) prank[i] = uint16_t(atomicAdd(gcnt + (pool[i].x >> 28), 1));
    lds_barrier();
    if (tid < VG && gcnt[tid] > 0) gbase[tid] = atomicAdd(a.cand_cnt + qid[tid], gcnt[tid]);
    lds_barrier();
    for (int i = tid; i < n; i += VNT) {
        const uint2 e = pool[i];
        const int g = int(e.x >> 28);
        const int pos = gbase[g] + int(prank[i]);
        const float ad = __builtin_bit_cast(float, e.y);
        if (pos < a.capq) {
            a.cand[int64_t(qid[g]) * a.capq + pos] = ckey(ad, int32_t(beg + (e.x & 0x0FFFFFFFu)));
            if (hsc[g] > 0.0f) atomicAdd(lhist + g * VNBH + min(VNBH - 1, max(0, int(ad * hsc[g]))), 1);
        }
    }
    lds_barrier();  // the pool is free again
    if (a.runthr) {
        // this flush's counts into the query's global histogram of candidate
        // distances (bin b: [b, b + 1) tau / VNBH); the first bin whose
        // cumulative count reaches kl bounds the kl-th smallest candidate
        // distance so far (counts read back may miss concurrent additions,
        // which only gives a larger, still valid, bound)
        for (int i = tid; i < VG * VNBH; i += VNT) {
            const int c = lhist[i];
            if (c > 0) atomicAdd(a.rhist + int64_t(qid[i / VNBH]) * fast_hist_bins_dev() + (i % VNBH), c);
        }
        const int lane = tid & 63;
        for (int g = tid >> 6; g < VG; g += VNT / 64) {  // one wavefront per slot
            if (!(hsc[g] > 0.0f) || gcnt[g] == 0) continue;  // wave-uniform
            const int q = qid[g];
            int c = lane < VNBH - 1 ? __hip_atomic_load(a.rhist + int64_t(q) * fast_hist_bins_dev() + lane,
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                    : 0;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int o = __shfl_up(c, off, 64);
                if (lane >= off) c += o;
            }
            const uint64_t reach = __ballot(lane < VNBH - 1 && c >= a.kl);
            if (reach != 0 && lane == 0) {
                const int b = __ffsll((unsigned long long)reach) - 1;
                const float R = (float(b + 1) + 0.01f) / hsc[g];
                atomicMin(reinterpret_cast<int *>(a.runthr) + q, __float_as_int(R));
            }
        }
    }
}

template <int M, int TT>
__global__ __launch_bounds__(VNT, (2 * VLay<M, TT>::total <= 160 * 1024 ? 4 : 2)) void fast_var_scan_kernel(FastArgs a) {
    using L = VLay<M, TT>;
    constexpr int RW = int(sizeof(FastItem) / 4);
    static_assert(RW <= 64, "a record is loaded by one wavefront");
    extern __shared__ __align__(16) unsigned char smem[];
    const float *lim_sh = reinterpret_cast<const float *>(smem + L::lim);
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
        if (tid == 0) atomicOr(a.err, 8);
        return;
    }
    float *taus = reinterpret_cast<float *>(smem + L::taus);
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
        if (tid < VG) {
            const int q = rec->q[tid];
            qid[tid] = q;
            taus[tid] = q >= 0 ? (a.tau ? a.tau[q] : rec->tau[tid]) : 0.0f;
        }
        if (tid < 3) pcnt[tid] = 0;
        __syncthreads();
        vtables<M, TT>(a, smem, l, taus);

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

        // per slot: the pass limit (registers) and the scale (FP16) or
        // offset (8-bit) of its sums (read from LDS each step)
        float lm[VG];
#pragma unroll
        for (int g = 0; g < VG; ++g) lm[g] = lim_sh[g];
        const uint32_t adj_at = uint32_t(TT == 1 ? L::mul : L::base);
        const uint8_t *lc = a.codes + beg * M;
        uint32_t cn[M / 4];
        vload_codes<M>(lc, min(tid, max(len - 1, 0)), cn);
        int fill = 0;  // staged candidates before this step (uniform)
        int rot = 0;   // append counter of this step
        for (int b0 = 0; b0 < len; b0 += VNT) {
            if (tid == 0) pcnt[rot == 2 ? 0 : rot + 1] = 0;
            const int v = b0 + tid;
            uint32_t cw[M / 4];
#pragma unroll
            for (int i = 0; i < M / 4; ++i) cw[i] = cn[i];
            if (b0 + VNT < len) vload_codes<M>(lc, min(v + VNT, len - 1), cn);
            float acc[VG];
            vadc<M, TT>(smem, cw, acc);
            uint32_t pm = 0u;
#pragma unroll
            for (int g4 = 0; g4 < VG / 4; ++g4) {
                const f32x4 adj = *reinterpret_cast<lds_f32x4 *>(uintptr_t(adj_at + 16u * g4));  // broadcast
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int g = 4 * g4 + j;
                    acc[g] = TT == 1 ? acc[g] * adj[j] : adj[j] + acc[g];
                    pm |= uint32_t(acc[g] <= lm[g]) << g;
                }
            }
            if (v >= len) pm = 0u;
            if (__ballot(pm != 0u) != 0) {  // wave-uniform
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
                if (lane == 0) slot = atomicAdd(pcnt + rot, int(tot));
                slot = fill + __builtin_amdgcn_readfirstlane(slot) + int(pre);
#pragma unroll
                for (int g = 0; g < VG; ++g) {
                    if (!((pm >> g) & 1u)) continue;
                    if (slot < VPOOL) {
                        pool[slot] = make_uint2(uint32_t(v) | (uint32_t(g) << 28), __builtin_bit_cast(uint32_t, acc[g]));
                    } else {  // pool full: straight to the global buffer
                        const int q = qid[g];
                        const int pos = atomicAdd(a.cand_cnt + q, 1);
                        if (pos < a.capq) a.cand[int64_t(q) * a.capq + pos] = ckey(acc[g], int32_t(beg + v));
                    }
                    ++slot;
                }
            }
            // three append counters rotate over the steps (as in the other
            // list-major scans): one LDS-only barrier per step
            lds_barrier();
            fill += pcnt[rot];
            rot = rot == 2 ? 0 : rot + 1;
            if (fill > VPOOL / 2) {  // uniform
                vflush<M, TT>(a, smem, beg, fill);
                fill = 0;
            }
            if (stage < 2) advance();
        }
        if (fill > 0) vflush<M, TT>(a, smem, beg, fill);
        while (stage < 2) {
            __syncthreads();
            advance();
        }
        __syncthreads();  // rec, qid and the pool are free; nrec is complete
        if (tid < RW) reinterpret_cast<uint32_t *>(rec)[tid] = reinterpret_cast<const uint32_t *>(nrec)[tid];
        if (tid == 0) *item_sh = *nidx_sh;
        __syncthreads();
    }
}

// Test hook: the approximate ADC of every vector of list l for queries
// 0..n-1 (out [len][16]; empty slots give whatever their zero residual gives).
template <int M, int TT>
__global__ __launch_bounds__(VNT) void fast_var_adc_kernel(FastArgs a, int l, int n, int64_t beg, int len,
                                                           float *out) {
    using L = VLay<M, TT>;
    extern __shared__ __align__(16) unsigned char smem[];
    int *qid = reinterpret_cast<int *>(smem + L::qid);
    float *taus = reinterpret_cast<float *>(smem + L::taus);
    const int tid = threadIdx.x;
    if (tid < VG) {
        qid[tid] = tid < n ? tid : -1;
        taus[tid] = __builtin_huge_valf();
    }
    __syncthreads();
    vtables<M, TT>(a, smem, l, taus);
    const float *mul_sh = reinterpret_cast<const float *>(smem + L::mul);
    const float *base_sh = reinterpret_cast<const float *>(smem + L::base);
    const uint8_t *lc = a.codes + beg * M;
    for (int v = tid; v < len; v += VNT) {
        uint32_t cw[M / 4];
        vload_codes<M>(lc, v, cw);
        float acc[VG];
        vadc<M, TT>(smem, cw, acc);
#pragma unroll
        for (int g = 0; g < VG; ++g) out[int64_t(v) * VG + g] = base_sh[g] + acc[g] * mul_sh[g];
    }
}

template <int M, int TT>
hipError_t set_lds() {
    static bool done = false;
    if (done) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&fast_var_scan_kernel<M, TT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(VLay<M, TT>::total));
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(&fast_var_adc_kernel<M, TT>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(VLay<M, TT>::total));
    done = e == hipSuccess;
    return e;
}

template <int M, int TT>
hipError_t launch_var(const FastArgs &a, int grid, hipStream_t s) {
    const hipError_t e = set_lds<M, TT>();
    if (e != hipSuccess) return e;
    fast_var_scan_kernel<M, TT><<<grid, VNT, VLay<M, TT>::total, s>>>(a);
    return hipGetLastError();
}

template <int M, int TT>
hipError_t launch_var_adc(const FastArgs &a, int l, int n, int64_t beg, int len, float *out, hipStream_t s) {
    const hipError_t e = set_lds<M, TT>();
    if (e != hipSuccess) return e;
    fast_var_adc_kernel<M, TT><<<1, VNT, VLay<M, TT>::total, s>>>(a, l, n, beg, len, out);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_fast_var_scan(const FastArgs &a, int m, int tables, int grid, hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    if (!a.pq || !a.pq_norm) return hipErrorInvalidValue;
    switch (m * 10 + tables) {
        case 81: return launch_var<8, 1>(a, grid, s);
        case 82: return launch_var<8, 2>(a, grid, s);
        case 161: return launch_var<16, 1>(a, grid, s);
        case 162: return launch_var<16, 2>(a, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_var_adc(const FastArgs &a, int m, int tables, int l, int n, int64_t beg, int len, float *out,
                               hipStream_t s) {
    if (n < 1 || n > VG || len < 0) return hipErrorInvalidValue;
    if (len == 0) return hipSuccess;
    switch (m * 10 + tables) {
        case 81: return launch_var_adc<8, 1>(a, l, n, beg, len, out, s);
        case 82: return launch_var_adc<8, 2>(a, l, n, beg, len, out, s);
        case 161: return launch_var_adc<16, 1>(a, l, n, beg, len, out, s);
        case 162: return launch_var_adc<16, 2>(a, l, n, beg, len, out, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mivf
