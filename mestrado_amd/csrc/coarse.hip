// Coarse assignment (top-w lists per query) with the reference's exact FP32
// arithmetic, computed in two steps:
//
// 1. coarse_approx_kernel: D'(q, c) = fl(fl(|q|^2) + |c|^2) - 2 q.c with the
//    dot products on the matrix cores in BF16 with FP32 accumulation
//    (v_mfma_f32_16x16x32_bf16).  Each operand is split x = hi + lo (two
//    BF16 terms, hi = bf16(x), lo = bf16(x - hi); centroids once per index,
//    queries by coarse_prep_kernel), and q.c is accumulated as
//    hi.hi + hi.lo + lo.hi: every BF16 x BF16 product is exact in FP32.
// 2. coarse_select_*_kernel: v' = an upper bound of the w-th smallest D', then
//    every centroid with D' <= v' + 2 M_q is re-evaluated with the exact
//    reference arithmetic (sequential FP32 dot, a = centroid with its
//    f32-accumulated norm, b = query with its f64-accumulated norm,
//    D = fl(fl(-2 dot) + fl(|q|^2 + |c|^2)), parallel/ivf_pq/ivf_assign.cpp:28
//    via knn_full) and the exact top-w by (distance, index) is kept.
//
// Why the recheck set holds the exact top-w.  Two BF16 terms keep 16
// significant bits, so x = hi + lo + r with |r| <= 2^-16 |x|; the dropped
// lo.lo and r terms bound the split product's error by 3 * 2^-16 |q_t c_t|,
// and FP32 accumulation of 3 * 128 products adds <= 384 * 2^-24 sum|q_t c_t|.
// With sum|q_t c_t| <= (|q|^2 + |c|^2) / 2, the approximate and the exact
// (sequential FP32) dot products differ by <= 3.3e-5 (|q|^2 + |c|^2); -2 dot
// doubles it and the norm terms add a few ulps, so |D' - D| <= M_q with
// M_q = 1.5e-4 (|q|^2 + max_c |c|^2) (over 2x headroom).  w centroids have
// D' <= v', hence D <= v' + M_q for them and the w-th exact distance is
// <= v' + M_q; a centroid of the exact top-w therefore has D' <= D + M_q <=
// v' + 2 M_q.  A centroid outside the set has D > v' + M_q >= the w-th exact
// distance, so it loses to all of the top-w, ties included.
#include <algorithm>
#include <atomic>

#include "select_dev.h"

namespace mivf {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CD = 128;       // dimension of the MFMA path
constexpr int CPB = 2 * CD + 8;  // padded LDS row of a query's hi and lo terms (bf16)
constexpr int NTC = 256;
constexpr int RCAP = 1024;    // recheck capacity per query

// Units of D': one 64-query tile x one column block of 64 TW centroids.  A
// persistent grid gives each workgroup a contiguous range of units (query
// tile major, balanced to within one unit); the query tile (BF16 hi and lo
// terms, one padded row per query) stays in LDS while the range stays in it.
// Wave wv owns centroids 16 TW wv .. 16 TW wv + 16 TW - 1 of each column
// block (TW 16-wide column tiles whose BF16 hi/lo rows are read from global
// memory one k block ahead, into the next unit too) and all four 16-query
// row tiles.  In a 32-wide k block, lane group h = l/16 supplies elements
// k0 + 8h .. k0 + 8h + 7 of both operands (the same k permutation on both
// sides of the sum).  TW = 2 keeps the accumulators at 32 registers.
constexpr int CQ = 64;   // queries per tile
constexpr int CTW = 2;   // column tiles per wave of the launched shape

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Query norms (f32 squares accumulated in f64 in element order, as
// launch_norm_f64) and the BF16 hi/lo split of the queries: 64 rows per
// block, loaded once through LDS.
__global__ __launch_bounds__(NTC) void coarse_prep_kernel(const float *__restrict__ queries, int nq,
                                                          double *__restrict__ qn, __bf16 *__restrict__ qhi,
                                                          __bf16 *__restrict__ qlo) {
    __shared__ float tile[64][CD + 1];
    const int tid = threadIdx.x;
    const int64_t r0 = int64_t(blockIdx.x) * 64;
    for (int e = tid; e < 64 * CD / 4; e += NTC) {
        const int rr = e / (CD / 4), cc = (e % (CD / 4)) * 4;
        const int64_t row = r0 + rr;
        if (row >= nq) continue;
        const float4 v = *reinterpret_cast<const float4 *>(queries + row * CD + cc);
        const float xs[4] = {v.x, v.y, v.z, v.w};
        __bf16 h[4], l[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            tile[rr][cc + u] = xs[u];
            h[u] = __bf16(xs[u]);
            l[u] = __bf16(xs[u] - float(h[u]));
        }
        *reinterpret_cast<uint2 *>(qhi + row * CD + cc) = *reinterpret_cast<const uint2 *>(h);
        *reinterpret_cast<uint2 *>(qlo + row * CD + cc) = *reinterpret_cast<const uint2 *>(l);
    }
    __syncthreads();
    if (tid < 64 && r0 + tid < nq) {
        double sum = 0.0;
        for (int t = 0; t < CD; ++t) {
            const float v = tile[tid][t];
            const float p = v * v;
            sum = sum + double(p);
        }
        qn[r0 + tid] = sum;
    }
}

// SPLIT: the query tile is split to BF16 here, from the FP32 queries, and the
// row norms of D' are FP32 sums over the tile (D' is approximate: an FP32
// norm in any order is within 128 * 2^-24 |q|^2 of the exact one, far inside
// the recheck margin); otherwise the terms and norms of coarse_prep_kernel
// are read.
template <int TW, bool SPLIT>
__global__ __launch_bounds__(NTC, 3) void coarse_approx_kernel(const __bf16 *__restrict__ chi,
                                                              const __bf16 *__restrict__ clo,
                                                              const float *__restrict__ cn, int nlist,
                                                              const float *__restrict__ queries,
                                                              const __bf16 *__restrict__ qhi,
                                                              const __bf16 *__restrict__ qlo,
                                                              const double *__restrict__ qn, int nq,
                                                              float *__restrict__ out, int64_t ldo,
                                                              float *__restrict__ gmin, int64_t ldg,
                                                              int *__restrict__ zero) {
    constexpr int CC = 64 * TW;          // centroids per column block
    constexpr int TS = 16 * TW + 4;      // staging row stride (floats)
    constexpr int RPS = 64 / (16 * TW);  // rows per store instruction
    extern __shared__ __align__(16) float lds[];
    __bf16 *Qs = reinterpret_cast<__bf16 *>(lds);  // [CQ][CPB]: hi terms, then lo terms
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 4, i16 = lane & 15;
    // this wave's epilogue staging tile, after the query tile; then the
    // tile's row norms (SPLIT)
    float *T = lds + CQ * CPB / 2 + wave * (16 * TS);
    float *qfs = lds + CQ * CPB / 2 + 4 * 16 * TS;
    if (zero && blockIdx.x == 0 && tid == 0) *zero = 0;  // the select's redo counter, for the next kernel
    const int ncb = (nlist + CC - 1) / CC;
    const int units = ((nq + CQ - 1) / CQ) * ncb;
    const int u0 = int(int64_t(units) * blockIdx.x / gridDim.x);
    const int u1 = int(int64_t(units) * (blockIdx.x + 1) / gridDim.x);
    // B operands (and centroid norms) of the next k block, one k block ahead
    bf16x8 bhn[TW], bln[TW];
    float cnn[TW];
    auto fetch = [&](int u, int k0) {
        const int cbn = u % ncb;
#pragma unroll
        for (int t = 0; t < TW; ++t) {
            const int c = cbn * CC + wave * (16 * TW) + 16 * t + i16;
            const int64_t off = int64_t(min(c, nlist - 1)) * CD + 8 * h + k0;
            bhn[t] = *reinterpret_cast<const bf16x8 *>(chi + off);
            bln[t] = *reinterpret_cast<const bf16x8 *>(clo + off);
            if (k0 == 0) cnn[t] = c < nlist ? cn[c] : 0.0f;
        }
    };
    if (u0 < u1) fetch(u0, 0);
    int qtile = -1;
    float qf[4][4];
    for (int u = u0; u < u1; ++u) {
        const int cb = u % ncb;
        if (u / ncb != qtile) {  // uniform: this unit's query tile
            if (qtile >= 0) __syncthreads();  // every wave is done with the previous tile
            qtile = u / ncb;
            const int64_t q0 = int64_t(qtile) * CQ;
            if constexpr (SPLIT) {
                // a row's 32 float4 are 32 consecutive lanes: one half-wave
                // sums its squares
                for (int e = tid; e < CQ * CD / 4; e += NTC) {
                    const int row = e / (CD / 4), col = (e % (CD / 4)) * 4;
                    float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (q0 + row < nq) x = *reinterpret_cast<const float4 *>(queries + (q0 + row) * CD + col);
                    const float xs[4] = {x.x, x.y, x.z, x.w};
                    __bf16 hh[4], ll[4];
                    float sq = 0.0f;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        hh[u] = __bf16(xs[u]);
                        ll[u] = __bf16(xs[u] - float(hh[u]));
                        sq = __builtin_fmaf(xs[u], xs[u], sq);
                    }
                    *reinterpret_cast<uint2 *>(Qs + row * CPB + col) = *reinterpret_cast<const uint2 *>(hh);
                    *reinterpret_cast<uint2 *>(Qs + row * CPB + CD + col) = *reinterpret_cast<const uint2 *>(ll);
#pragma unroll
                    for (int off = 16; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 32);
                    if ((e & 31) == 0) qfs[row] = sq;
                }
                __syncthreads();
#pragma unroll
                for (int qt = 0; qt < 4; ++qt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) qf[qt][r] = qfs[16 * qt + 4 * h + r];
            } else {
                for (int e = tid; e < CQ * 2 * CD / 8; e += NTC) {
                    const int row = e / (2 * CD / 8), part = (e % (2 * CD / 8)) * 8;  // part < CD: hi, else lo
                    uint4 x = make_uint4(0u, 0u, 0u, 0u);
                    if (q0 + row < nq)
                        x = *reinterpret_cast<const uint4 *>((part < CD ? qhi : qlo) + (q0 + row) * CD + (part % CD));
                    *reinterpret_cast<uint4 *>(Qs + row * CPB + part) = x;
                }
                // the norms of this lane's 16 query rows
#pragma unroll
                for (int qt = 0; qt < 4; ++qt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) qf[qt][r] = float(qn[min(q0 + 16 * qt + 4 * h + r, int64_t(nq) - 1)]);
            }
            __syncthreads();
        }
        const int64_t q0 = int64_t(qtile) * CQ;
        const int c0 = cb * CC + wave * (16 * TW);
        float cnc[TW];
#pragma unroll
        for (int t = 0; t < TW; ++t) cnc[t] = cnn[t];
        f32x4 acc[4][TW];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int t = 0; t < TW; ++t) acc[a][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int k0 = 0; k0 < CD; k0 += 32) {
            bf16x8 bh[TW], bl[TW];
#pragma unroll
            for (int t = 0; t < TW; ++t) {
                bh[t] = bhn[t];
                bl[t] = bln[t];
            }
            // the next k block: this unit's next one, or the next unit's first
            if (k0 + 32 < CD)
                fetch(u, k0 + 32);
            else if (u + 1 < u1)
                fetch(u + 1, 0);
#pragma unroll
            for (int qt = 0; qt < 4; ++qt) {
                const __bf16 *qr = Qs + (16 * qt + i16) * CPB + k0 + 8 * h;
                const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(qr);
                const bf16x8 al = *reinterpret_cast<const bf16x8 *>(qr + CD);
#pragma unroll
                for (int t = 0; t < TW; ++t) {
                    acc[qt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[t], acc[qt][t], 0, 0, 0);
                    acc[qt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[t], acc[qt][t], 0, 0, 0);
                    acc[qt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[t], acc[qt][t], 0, 0, 0);
                }
            }
        }
        // accumulator r of lane l: query row 16 qt + 4h + r, centroid column
        // 16 t + l % 16.  Each 16 x 16 TW block of the wave goes through LDS so
        // that a store writes whole row pieces (16 TW centroids, 64 TW
        // contiguous bytes; with TW = 2 the two halves of the wave store two
        // rows).  Row stride TS = 16 TW + 4 floats puts the four rows a write
        // touches (h = 0..3) in distinct banks.
        const int sc = lane % (16 * TW), sr = lane / (16 * TW);  // this lane's column and row within a store
        const int cl = c0 + sc;
#pragma unroll
        for (int qt = 0; qt < 4; ++qt) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
#pragma unroll
                for (int t = 0; t < TW; ++t)
                    T[(4 * h + r) * TS + 16 * t + i16] = (qf[qt][r] + cnc[t]) - 2.0f * acc[qt][t][r];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // the tile's values are read from LDS together, then stored
            // (a whole tile without per-row checks)
            float ov[16 / RPS];
#pragma unroll
            for (int i = 0; i < 16 / RPS; ++i) ov[i] = T[(RPS * i + sr) * TS + sc];
            const int64_t qr = q0 + 16 * qt + sr;
            if (cl < nlist) {
                if (q0 + 16 * qt + 16 <= nq) {  // uniform
#pragma unroll
                    for (int i = 0; i < 16 / RPS; ++i) out[(qr + RPS * i) * ldo + cl] = ov[i];
                } else {
#pragma unroll
                    for (int i = 0; i < 16 / RPS; ++i)
                        if (qr + RPS * i < nq) out[(qr + RPS * i) * ldo + cl] = ov[i];
                }
            }
            if (gmin && lane < 16 * TW) {
                // the minimum of D' over the 16 centroids of one column tile
                // (group (c0 + 16 g) / 16) for one query row
                const int row = lane / TW, g = lane % TW;
                const int cg = c0 + 16 * g;
                const int64_t q = q0 + 16 * qt + row;
                if (q < nq && cg < nlist) {
                    const float4 *tr = reinterpret_cast<const float4 *>(T + row * TS + 16 * g);
                    float mv = __builtin_huge_valf();
#pragma unroll
                    for (int j4 = 0; j4 < 4; ++j4) {
                        const float4 x = tr[j4];
                        const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            if (cg + 4 * j4 + v < nlist) mv = fminf(mv, xs[v]);
                    }
                    gmin[q * ldg + cg / 16] = mv;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

// BF16 hi/lo split of the centroids (once per index).
__global__ void coarse_split_kernel(const float *__restrict__ coarse, int64_t n, __bf16 *__restrict__ hi,
                                    __bf16 *__restrict__ lo) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = coarse[i];
    const __bf16 h = __bf16(x);
    hi[i] = h;
    lo[i] = __bf16(x - float(h));
}

// Exact reference distance of centroid c to the query in LDS.
__device__ inline float exact_dist(const float *__restrict__ crow, const float *qv, float an, double bn) {
    float acc = 0.0f;
    // the row in four batches of 32 values, each batch's loads issued together
#pragma unroll 1
    for (int t0 = 0; t0 < CD; t0 += 32) {
        float4 c4[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) c4[j] = *reinterpret_cast<const float4 *>(crow + t0 + 4 * j);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float cc[4] = {c4[j].x, c4[j].y, c4[j].z, c4[j].w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float p = cc[u] * qv[t0 + 4 * j + u];
                acc = acc + p;
            }
        }
    }
    const float c = float(bn + double(an));
    const float m2 = -2.0f * acc;
    return m2 + c;
}

__global__ __launch_bounds__(NTC) void coarse_select_kernel(const float *__restrict__ coarse,
                                                            const float *__restrict__ cn, float maxcn, int nlist,
                                                            const float *__restrict__ queries,
                                                            const double *__restrict__ qn,
                                                            const float *__restrict__ approx, int64_t lda, int w,
                                                            int32_t *__restrict__ probes,
                                                            float *__restrict__ pdis,
                                                            const int32_t *__restrict__ only,
                                                            const int *__restrict__ only_cnt) {
    extern __shared__ __align__(16) unsigned char smem[];
    float *row = reinterpret_cast<float *>(smem);                        // [nlist] D', then exact D (fallback)
    uint64_t *ck = reinterpret_cast<uint64_t *>(smem + ((size_t(nlist) * 4 + 15) & ~size_t(15)));  // [RCAP]
    __shared__ SelectShared sh;
    __shared__ uint64_t buf[RCAP];
    __shared__ float qv[CD];
    __shared__ int ncand, cnt, eqt;
    const int tid = threadIdx.x;
    // without `only`: query blockIdx.x; with it: list entries blockIdx.x,
    // blockIdx.x + gridDim.x, ...
    for (int it = int(blockIdx.x);; it += int(gridDim.x)) {
        if (only ? it >= *only_cnt : it != int(blockIdx.x)) break;  // uniform
        const int64_t q = only ? only[it] : int64_t(it);
        const double bn = qn[q];
        if (tid < CD) qv[tid] = queries[q * CD + tid];
        if (tid == 0) ncand = 0;
        for (int c = tid; c < nlist; c += NTC) row[c] = approx[q * lda + c];
        // An upper bound V of the w-th smallest D': a 256-bin histogram over
        // [min, max], then V = the largest D' in the bins up to the one where the
        // count reaches w (at least w values are <= V).
        float mn = __builtin_huge_valf(), mx = -__builtin_huge_valf();
        for (int c = tid; c < nlist; c += NTC) {
            mn = fminf(mn, row[c]);
            mx = fmaxf(mx, row[c]);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, off));
            mx = fmaxf(mx, __shfl_xor(mx, off));
        }
        __shared__ float wmn[NTC / 64], wmx[NTC / 64];
        __shared__ uint32_t hist[256];
        __shared__ int sdig;
        __shared__ uint32_t vmax;
        if ((tid & 63) == 0) {
            wmn[tid >> 6] = mn;
            wmx[tid >> 6] = mx;
        }
        hist[tid] = 0u;
        if (tid == 0) vmax = 0u;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NTC / 64; ++i) {
            mn = fminf(mn, wmn[i]);
            mx = fmaxf(mx, wmx[i]);
        }
        const float scale = mx > mn ? 256.0f / (mx - mn) : 0.0f;
        auto bin = [&](float v) { return min(255, max(0, int((v - mn) * scale))); };
        for (int c = tid; c < nlist; c += NTC) atomicAdd(&hist[bin(row[c])], 1u);
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
            const uint32_t need = uint32_t(min(w, nlist));
            if (excl < need && incl >= need) {
                uint32_t c = excl;
                int dg;
                if (c + h0 >= need) dg = 0;
                else if (c + h0 + h1 >= need) { c += h0; dg = 1; }
                else if (c + h0 + h1 + h2 >= need) { c += h0 + h1; dg = 2; }
                else { c += h0 + h1 + h2; dg = 3; }
                sdig = 4 * tid + dg;
            }
        }
        __syncthreads();
        const int bmax = sdig;
        uint32_t mine = 0u;
        for (int c = tid; c < nlist; c += NTC)
            if (bin(row[c]) <= bmax) mine = max(mine, fkey(row[c]));
        atomicMax(&vmax, mine);
        __syncthreads();
        const float vw = key_dist(uint64_t(vmax) << 32);
        auto akey = [&](int i) -> uint64_t { return ckey(row[i], i); };
        const float lim = vw + 3.0e-4f * float(bn + double(maxcn));
        // recheck set
        for (int i = tid; i < nlist; i += NTC) {
            if (row[i] <= lim) {
                const int p = atomicAdd(&ncand, 1);
                if (p < RCAP) ck[p] = uint64_t(i);
            }
        }
        __syncthreads();
        const int n = ncand;
        SelectResult r;
        int nk;
        if (n <= RCAP) {
            for (int i = tid; i < n; i += NTC) {
                const int c = int(ck[i]);
                ck[i] = ckey(exact_dist(coarse + int64_t(c) * CD, qv, cn[c], bn), c);
            }
            __syncthreads();
            auto kf = [&](int i) { return ck[i]; };
            nk = n;
            r = block_select<NTC>(kf, nk, w, sh);
        } else {
            // recheck set too large: exact distances of every centroid
            __syncthreads();
            for (int c = tid; c < nlist; c += NTC) row[c] = exact_dist(coarse + int64_t(c) * CD, qv, cn[c], bn);
            __syncthreads();
            nk = nlist;
            r = block_select<NTC>(akey, nk, w, sh);
        }
        // collect the w best into buf and sort
        if (tid == 0) {
            cnt = 0;
            eqt = 0;
        }
        __syncthreads();
        for (int i = tid; i < nk; i += NTC) {
            const uint64_t k = (n <= RCAP) ? ck[i] : akey(i);
            bool take = k < r.thr;
            if (!take && r.eq > 0 && k == r.thr) take = atomicAdd(&eqt, 1) < r.eq;
            if (take) {
                const int p = atomicAdd(&cnt, 1);
                if (p < RCAP) buf[p] = k;
            }
        }
        __syncthreads();
        const int c = min(cnt, RCAP);
        const int np = next_pow2(c > 0 ? c : 1);
        for (int i = c + tid; i < np; i += NTC) buf[i] = kNoKey;
        __syncthreads();
        block_bitonic_sort<NTC>(buf, np);
        for (int i = tid; i < w; i += NTC) {
            probes[q * w + i] = i < c ? key_id(buf[i]) : -1;
            pdis[q * w + i] = i < c ? key_dist(buf[i]) : __builtin_huge_valf();
        }
        __syncthreads();  // the shared buffers are reused by the next query
    }
}

// One wavefront per query (nlist <= 64 * NPL, w <= 64): lane l holds the D'
// of centroids l, l + 64, ...; the w-th smallest of the 64 lane minima is an
// upper bound V of the w-th smallest D' (the minima are w distinct values).
// The recheck set {D' <= V + 2 M_q} is compacted by ballots, evaluated with
// the exact arithmetic, and its w best found by rank counting (keys are
// distinct).  A query whose recheck set exceeds WCAP (or is short of w, which
// only non-finite distances cause) goes to `redo` for the workgroup kernel.
constexpr int NPL = 64;
constexpr int WCAP = 256;

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(NTC) void coarse_select_wave_kernel(const float *__restrict__ coarse,
                                                                 const float *__restrict__ cn, float maxcn,
                                                                 int nlist, const float *__restrict__ queries,
                                                                 const double *__restrict__ qn, int nq,
                                                                 const float *__restrict__ approx, int64_t lda,
                                                                 int w, int32_t *__restrict__ probes,
                                                                 float *__restrict__ pdis,
                                                                 int32_t *__restrict__ redo, int *__restrict__ redo_cnt) {
    __shared__ uint32_t cidx_all[NTC / 64][WCAP];
    __shared__ uint64_t keys_all[NTC / 64][WCAP];
    __shared__ float qv_all[NTC / 64][CD];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t q = int64_t(blockIdx.x) * (NTC / 64) + wv;
    if (q >= nq) return;  // wave-uniform; no workgroup barriers below
    uint32_t *cidx = cidx_all[wv];
    uint64_t *keys = keys_all[wv];
    float *qv = qv_all[wv];
    const double bn = qn[q];
    qv[lane] = queries[q * CD + lane];
    qv[lane + 64] = queries[q * CD + lane + 64];
    float v[NPL];
    float mn = __builtin_huge_valf();
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
        const int c = i * 64 + lane;
        v[i] = c < nlist ? approx[q * lda + c] : __builtin_huge_valf();
        mn = fminf(mn, v[i]);
    }
    // bitonic sort of the 64 lane minima across the wavefront (ascending)
    float x = mn;
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const float y = __shfl_xor(x, j);
            const bool up = (lane & k) == 0;
            const bool lower = (lane & j) == 0;
            x = (lower == up) ? fminf(x, y) : fmaxf(x, y);
        }
    }
    const float vw = __shfl(x, w - 1);
    const float lim = vw + 3.0e-4f * float(bn + double(maxcn));
    int n = 0;
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
        const bool pass = v[i] <= lim;  // +inf padding never passes a finite limit
        const uint64_t bal = __ballot(pass);
        const int pos = n + int(__builtin_amdgcn_mbcnt_hi(uint32_t(bal >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(bal), 0u)));
        if (pass && pos < WCAP) cidx[pos] = uint32_t(i * 64 + lane);
        n += __popcll(bal);
    }
    if (n > WCAP || n < w) {
        if (lane == 0) redo[atomicAdd(redo_cnt, 1)] = int32_t(q);
        return;
    }
    wave_sync();
    for (int j = lane; j < n; j += 64) {
        const int c = int(cidx[j]);
        keys[j] = ckey(exact_dist(coarse + int64_t(c) * CD, qv, cn[c], bn), c);
    }
    wave_sync();
    for (int j = lane; j < n; j += 64) {
        const uint64_t k = keys[j];
        int rank = 0;
        for (int i = 0; i < n; ++i) rank += keys[i] < k;
        if (rank < w) {
            probes[q * w + rank] = key_id(k);
            pdis[q * w + rank] = key_dist(k);
        }
    }
}

// One wavefront per query from the 16-centroid group minima of D' (written
// by coarse_approx_kernel): the w-th smallest group minimum V bounds the w-th
// smallest D' from above (the minima are w distinct centroids), found by a
// bitwise radix search over ballot counts of their order keys; a centroid of
// the recheck set {D' <= V + 2 M_q} lies in a group whose minimum is within
// the limit, so only those groups' D' are read (16 lanes per group).  Lane l
// holds the minima of groups l + 64 j, j < GM.  A query whose passing groups
// exceed GCAP, or whose recheck set exceeds WCAP or is short of w, goes to
// `redo`.
constexpr int GCAP = 256;

template <int GM>
__global__ __launch_bounds__(NTC) void coarse_select_group_kernel(const float *__restrict__ coarse,
                                                                  const float *__restrict__ cn, float maxcn,
                                                                  int nlist, const float *__restrict__ queries,
                                                                  const double *__restrict__ qn, int nq,
                                                                  const float *__restrict__ approx, int64_t lda,
                                                                  const float *__restrict__ gmin, int64_t ldg,
                                                                  int w, int32_t *__restrict__ probes,
                                                                  float *__restrict__ pdis,
                                                                  int32_t *__restrict__ redo,
                                                                  int *__restrict__ redo_cnt,
                                                                  double *__restrict__ qn_out) {
    __shared__ uint32_t cidx_all[NTC / 64][WCAP];
    __shared__ uint64_t keys_all[NTC / 64][WCAP];
    __shared__ float qv_all[NTC / 64][CD];
    __shared__ uint16_t gl_all[NTC / 64][GCAP];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t q = int64_t(blockIdx.x) * (NTC / 64) + wv;
    if (q >= nq) return;  // wave-uniform; no workgroup barriers below
    uint32_t *cidx = cidx_all[wv];
    uint64_t *keys = keys_all[wv];
    float *qv = qv_all[wv];
    uint16_t *gl = gl_all[wv];
    const int ng = (nlist + 15) / 16;
    const float *gq = gmin + q * ldg;
    uint32_t gk[GM];
#pragma unroll
    for (int j = 0; j < GM; ++j) {
        const int g = lane + 64 * j;
        gk[j] = fkey(g < ng ? gq[g] : __builtin_huge_valf());
    }
    qv[lane] = queries[q * CD + lane];
    qv[lane + 64] = queries[q * CD + lane + 64];
    double bn;
    if (qn_out) {
        // the query's norm (f32 squares summed in f64 in element order, as
        // launch_norm_f64), for this kernel and the later ones
        wave_sync();
        double sum = 0.0;
        if (lane == 0) {
            for (int t = 0; t < CD; ++t) {
                const float p = qv[t] * qv[t];
                sum = sum + double(p);
            }
            qn_out[q] = sum;
        }
        const uint64_t bits = __builtin_bit_cast(uint64_t, sum);
        const uint32_t blo = uint32_t(__shfl(int(uint32_t(bits)), 0));
        const uint32_t bhi = uint32_t(__shfl(int(uint32_t(bits >> 32)), 0));
        bn = __builtin_bit_cast(double, (uint64_t(bhi) << 32) | blo);
    } else {
        bn = qn[q];
    }
    // An upper bound of the w-th smallest group minimum: MSB-first search for
    // the smallest key with at least w minima at or below it (integer order
    // = float order), from the first bit below those all minima share, over
    // 16 bits; the bound is the largest key of the bucket found.
    uint32_t kmn = 0xFFFFFFFFu, kmx = 0u;
#pragma unroll
    for (int j = 0; j < GM; ++j)
        if (lane + 64 * j < ng) {
            kmn = min(kmn, gk[j]);
            kmx = max(kmx, gk[j]);
        }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        kmn = min(kmn, uint32_t(__shfl_xor(int(kmn), off)));
        kmx = max(kmx, uint32_t(__shfl_xor(int(kmx), off)));
    }
    const int cp = kmn == kmx ? 32 : __clz(int(kmn ^ kmx));  // shared leading bits
    uint32_t prefix = cp == 32 ? kmn : (cp == 0 ? 0u : (kmn >> (32 - cp)) << (32 - cp));
    const int top = 31 - cp, bot = max(0, top - 15);
    int need = w;
    for (int bit = top; bit >= bot; --bit) {
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < GM; ++j) cnt += __popcll(__ballot(((gk[j] ^ prefix) >> bit) == 0u));
        if (cnt < need) {
            need -= cnt;
            prefix |= 1u << bit;
        }
    }
    if (bot > 0) prefix |= (1u << bot) - 1u;
    const float vw = key_dist(uint64_t(prefix) << 32);
    const float lim = vw + 3.0e-4f * float(bn + double(maxcn));
    const uint32_t klim = fkey(lim);
    int ngp = 0;
#pragma unroll
    for (int j = 0; j < GM; ++j) {
        const int g = lane + 64 * j;
        const bool pass = g < ng && gk[j] <= klim;
        const uint64_t bal = __ballot(pass);
        const int pos = ngp + int(__builtin_amdgcn_mbcnt_hi(uint32_t(bal >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(bal), 0u)));
        if (pass && pos < GCAP) gl[pos] = uint16_t(g);
        ngp += __popcll(bal);
    }
    if (ngp > GCAP || !(vw < __builtin_huge_valf())) {
        if (lane == 0) redo[atomicAdd(redo_cnt, 1)] = int32_t(q);
        return;
    }
    wave_sync();
    // D' of the passing groups, four groups per load round, twelve rounds in flight
    constexpr int RB = 12;
    const float *row = approx + q * lda;
    int n = 0;
    for (int g0 = 0; g0 < ngp; g0 += 4 * RB) {
        float v[RB];
        int c[RB];
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            const int gi = g0 + 4 * u + (lane >> 4);
            c[u] = gi < ngp ? 16 * int(gl[gi]) + (lane & 15) : nlist;
            v[u] = c[u] < nlist ? row[c[u]] : __builtin_huge_valf();
        }
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            const bool pass = c[u] < nlist && v[u] <= lim;
            const uint64_t bal = __ballot(pass);
            const int pos = n + int(__builtin_amdgcn_mbcnt_hi(uint32_t(bal >> 32),
                                                              __builtin_amdgcn_mbcnt_lo(uint32_t(bal), 0u)));
            if (pass && pos < WCAP) cidx[pos] = uint32_t(c[u]);
            n += __popcll(bal);
        }
    }
    if (n > WCAP || n < w) {
        if (lane == 0) redo[atomicAdd(redo_cnt, 1)] = int32_t(q);
        return;
    }
    wave_sync();
    for (int j = lane; j < n; j += 64) {
        const int ci = int(cidx[j]);
        keys[j] = ckey(exact_dist(coarse + int64_t(ci) * CD, qv, cn[ci], bn), ci);
    }
    wave_sync();
    for (int j = lane; j < n; j += 64) {
        const uint64_t k = keys[j];
        int rank = 0;
        for (int i = 0; i < n; ++i) rank += keys[i] < k;
        if (rank < w) {
            probes[q * w + rank] = key_id(k);
            pdis[q * w + rank] = key_dist(k);
        }
    }
}

// The same per wavefront for any nlist up to the matrix-core limit and w up to
// 128: the approximate distances stream from HBM twice instead of sitting in
// registers.  Lane l owns the GPL groups of centroids {64 (GPL i + j) + l}
// (j < GPL; GPL >= w / 16, so the bound sits near the quarter quantile of
// the group minima rather than at their maximum).  The w-th smallest of the
// 64 GPL group minima bounds the w-th smallest D' from above (the minima are
// distinct centroids); it is found by a bitwise radix search over ballot
// counts of the minima's order keys.  The second pass compacts the recheck
// set {D' <= V + 2 M_q} by ballots; its exact distances are ranked as in the
// register kernel.  A recheck set above SCAP_W (or short of w) sends the
// query to `redo`.
constexpr int SCAP_W = 512;

template <int GPL>
__global__ __launch_bounds__(NTC) void coarse_select_stream_kernel(const float *__restrict__ coarse,
                                                                   const float *__restrict__ cn, float maxcn,
                                                                   int nlist, const float *__restrict__ queries,
                                                                   const double *__restrict__ qn, int nq,
                                                                   const float *__restrict__ approx, int64_t lda,
                                                                   int w, int32_t *__restrict__ probes,
                                                                   float *__restrict__ pdis,
                                                                   int32_t *__restrict__ redo,
                                                                   int *__restrict__ redo_cnt) {
    __shared__ uint32_t cidx_all[NTC / 64][SCAP_W];
    __shared__ uint64_t keys_all[NTC / 64][SCAP_W];
    __shared__ float qv_all[NTC / 64][CD];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t q = int64_t(blockIdx.x) * (NTC / 64) + wv;
    if (q >= nq) return;  // wave-uniform; no workgroup barriers below
    uint32_t *cidx = cidx_all[wv];
    uint64_t *keys = keys_all[wv];
    float *qv = qv_all[wv];
    const double bn = qn[q];
    qv[lane] = queries[q * CD + lane];
    qv[lane + 64] = queries[q * CD + lane + 64];
    const float *row = approx + q * lda;
    const int nblk = (nlist + 63) / 64;  // blocks of 64 centroids; block b belongs to group set b % GPL
    constexpr int UL = 16;               // blocks loaded together (loads in flight per lane)
    static_assert(UL % GPL == 0, "a load batch covers whole group sets");
    float gm[GPL];
#pragma unroll
    for (int j = 0; j < GPL; ++j) gm[j] = __builtin_huge_valf();
    for (int b0 = 0; b0 < nblk; b0 += UL) {
        float v[UL];
#pragma unroll
        for (int u = 0; u < UL; ++u) {
            const int c = (b0 + u) * 64 + lane;
            v[u] = c < nlist ? row[c] : __builtin_huge_valf();
        }
#pragma unroll
        for (int u = 0; u < UL; ++u) gm[u % GPL] = fminf(gm[u % GPL], v[u]);
    }
    // the w-th smallest group minimum: MSB-first search for the smallest key
    // with at least w minima at or below it (order keys: integer order =
    // float order)
    uint32_t gk[GPL];
#pragma unroll
    for (int j = 0; j < GPL; ++j) gk[j] = fkey(gm[j]);
    uint32_t prefix = 0u;
    int need = w;
    for (int bit = 31; bit >= 0; --bit) {
        // minima that agree with prefix above `bit` and have 0 at `bit`
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < GPL; ++j) cnt += __popcll(__ballot(((gk[j] ^ prefix) >> bit) == 0u));
        if (cnt < need) {
            need -= cnt;
            prefix |= 1u << bit;
        }
    }
    const float vw = key_dist(uint64_t(prefix) << 32);
    const float lim = vw + 3.0e-4f * float(bn + double(maxcn));
    int n = 0;
    for (int b0 = 0; b0 < nblk; b0 += UL) {
        float v[UL];
#pragma unroll
        for (int u = 0; u < UL; ++u) {
            const int c = (b0 + u) * 64 + lane;
            v[u] = c < nlist ? row[c] : __builtin_huge_valf();
        }
#pragma unroll
        for (int u = 0; u < UL; ++u) {
            const int c = (b0 + u) * 64 + lane;
            const bool pass = c < nlist && v[u] <= lim;
            const uint64_t bal = __ballot(pass);
            const int pos = n + int(__builtin_amdgcn_mbcnt_hi(uint32_t(bal >> 32),
                                                              __builtin_amdgcn_mbcnt_lo(uint32_t(bal), 0u)));
            if (pass && pos < SCAP_W) cidx[pos] = uint32_t(c);
            n += __popcll(bal);
        }
    }
    if (n > SCAP_W || n < w || !(vw < __builtin_huge_valf())) {
        if (lane == 0) redo[atomicAdd(redo_cnt, 1)] = int32_t(q);
        return;
    }
    wave_sync();
    for (int j = lane; j < n; j += 64) {
        const int c = int(cidx[j]);
        keys[j] = ckey(exact_dist(coarse + int64_t(c) * CD, qv, cn[c], bn), c);
    }
    wave_sync();
    for (int j = lane; j < n; j += 64) {
        const uint64_t k = keys[j];
        int rank = 0;
        for (int i = 0; i < n; ++i) rank += keys[i] < k;
        if (rank < w) {
            probes[q * w + rank] = key_id(k);
            pdis[q * w + rank] = key_dist(k);
        }
    }
}

}  // namespace

// The workgroup select holds a query's nlist approximate distances in LDS
// (plus RCAP keys): up to 32768 lists, 136 KiB of the 160 KiB.
constexpr int kMaxMfmaLists = 32768;

bool coarse_mfma_ok(int d, int nlist, int w) {
    return d == CD && nlist >= 1 && nlist <= kMaxMfmaLists && w >= 1 && w <= RCAP;
}

size_t coarse_select_lds(int nlist) { return ((size_t(nlist) * 4 + 15) & ~size_t(15)) + RCAP * 8; }

hipError_t launch_coarse_split(const float *coarse, int64_t n, void *hi, void *lo, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    coarse_split_kernel<<<unsigned((n + 255) / 256), 256, 0, s>>>(coarse, n, static_cast<__bf16 *>(hi),
                                                                  static_cast<__bf16 *>(lo));
    return hipGetLastError();
}

hipError_t launch_coarse_prep(const float *queries, int nq, double *qn, void *qhi, void *qlo, hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    coarse_prep_kernel<<<unsigned((nq + 63) / 64), NTC, 0, s>>>(queries, nq, qn, static_cast<__bf16 *>(qhi),
                                                                static_cast<__bf16 *>(qlo));
    return hipGetLastError();
}

hipError_t launch_coarse_approx(const void *chi, const void *clo, const float *cn, int nlist, const float *queries,
                                const void *qhi, const void *qlo, const double *qn, int nq, float *approx,
                                int64_t lda, float *gmin, int64_t ldg, int *zero, hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    constexpr int CC = 64 * CTW;  // centroids per column block
    const int64_t units = int64_t((nq + CQ - 1) / CQ) * ((nlist + CC - 1) / CC);
    if (units >= (int64_t(1) << 31)) return hipErrorInvalidValue;
    // three workgroups per CU (their LDS: query tile + staging tiles)
    static std::atomic<int> cus[64] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    int ncu = cus[dev].load(std::memory_order_relaxed);
    if (!ncu) {
        e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        cus[dev].store(ncu, std::memory_order_relaxed);
    }
    const unsigned grid = unsigned(std::min<int64_t>(units, int64_t(3) * std::max(ncu, 1)));
    const size_t lds = size_t(CQ) * CPB * 2 + 4 * 16 * (16 * CTW + 4) * 4 + CQ * 4;
    if (queries)
        coarse_approx_kernel<CTW, true><<<grid, NTC, lds, s>>>(
            static_cast<const __bf16 *>(chi), static_cast<const __bf16 *>(clo), cn, nlist, queries, nullptr,
            nullptr, nullptr, nq, approx, lda, gmin, ldg, zero);
    else
        coarse_approx_kernel<CTW, false><<<grid, NTC, lds, s>>>(
            static_cast<const __bf16 *>(chi), static_cast<const __bf16 *>(clo), cn, nlist, nullptr,
            static_cast<const __bf16 *>(qhi), static_cast<const __bf16 *>(qlo), qn, nq, approx, lda, gmin, ldg,
            zero);
    return hipGetLastError();
}

bool coarse_group_select_ok(int nlist, int w) { return nlist >= 64 && w <= 64 && (nlist + 15) / 16 >= w; }

hipError_t launch_coarse_select(const float *coarse, const float *cn, float maxcn, int nlist, const float *queries,
                                const double *qn, int nq, int w, const float *approx, int64_t lda,
                                const float *gmin, int64_t ldg, int32_t *probes, float *pdis, int32_t *redo,
                                int *redo_cnt, double *qn_out, hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    if (nlist > kMaxMfmaLists) return hipErrorInvalidValue;
    // the norms come from the group select only
    if (qn_out && !(gmin && redo && redo_cnt && coarse_group_select_ok(nlist, w))) return hipErrorInvalidValue;
    static size_t lds_set = 64 * 1024;  // dynamic LDS allowed so far (64 KiB needs no attribute)
    const size_t lds = coarse_select_lds(nlist);
    if (lds > lds_set) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&coarse_select_kernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
        lds_set = lds;
    }
    const unsigned wgrid = unsigned((nq + NTC / 64 - 1) / (NTC / 64));
    if (nlist >= 64 && w <= 128 && redo && redo_cnt) {
#define MIVF_STREAM_SELECT(G)                                                                                   \
    coarse_select_stream_kernel<G><<<wgrid, NTC, 0, s>>>(coarse, cn, maxcn, nlist, queries, qn, nq, approx, lda, w, \
                                                         probes, pdis, redo, redo_cnt)
        const int ngm = (nlist + 15) / 16 <= 256 ? 4 : (nlist + 15) / 16 <= 512 ? 8 : (nlist + 15) / 16 <= 1024 ? 16 : 32;
#define MIVF_GROUP_SELECT(G)                                                                                    \
    coarse_select_group_kernel<G><<<wgrid, NTC, 0, s>>>(coarse, cn, maxcn, nlist, queries, qn, nq, approx, lda,  \
                                                        gmin, ldg, w, probes, pdis, redo, redo_cnt, qn_out)
        const bool grp = gmin && coarse_group_select_ok(nlist, w);
        if (grp && ngm == 4)
            MIVF_GROUP_SELECT(4);
        else if (grp && ngm == 8)
            MIVF_GROUP_SELECT(8);
        else if (grp && ngm == 16)
            MIVF_GROUP_SELECT(16);
        else if (grp)
            MIVF_GROUP_SELECT(32);
#undef MIVF_GROUP_SELECT
        else if (nlist <= 64 * NPL && w <= 32)
            coarse_select_wave_kernel<<<wgrid, NTC, 0, s>>>(coarse, cn, maxcn, nlist, queries, qn, nq, approx, lda,
                                                            w, probes, pdis, redo, redo_cnt);
        else if (w <= 16)
            MIVF_STREAM_SELECT(1);
        else if (w <= 32)
            MIVF_STREAM_SELECT(2);
        else if (w <= 64)
            MIVF_STREAM_SELECT(4);
        else
            MIVF_STREAM_SELECT(8);
#undef MIVF_STREAM_SELECT
        // the few queries sent back: a small grid walks the list
        coarse_select_kernel<<<unsigned(std::min(nq, 256)), NTC, coarse_select_lds(nlist), s>>>(
            coarse, cn, maxcn, nlist, queries, qn, approx, lda, w, probes, pdis, redo, redo_cnt);
    } else {
        coarse_select_kernel<<<unsigned(nq), NTC, coarse_select_lds(nlist), s>>>(
            coarse, cn, maxcn, nlist, queries, qn, approx, lda, w, probes, pdis, nullptr, nullptr);
    }
    return hipGetLastError();
}

}  // namespace mivf
