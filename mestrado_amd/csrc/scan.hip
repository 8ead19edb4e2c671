// Fused IVF-PQ search kernel: for one query and a group of its probed lists,
//   residual r = fl32(q - c_list)                 (parallel/ivf_pq/ivf_assign.cpp:72-78)
//   table T[s][j] = dist(a = r_s, b = P[s][j])     (parallel/ivf_pq/ivf_search.cpp:206-209)
//   ADC  d = 0f; d = fl32(d + T[s][code_s]), s ascending   (ivf_search.cpp:451-468)
//   running exact top-k by (distance, id)          (parallel2/ivf_pq/topk.cu:26-32)
//
// Workgroup = 512 threads (8 wavefronts).  Thread (j, h) keeps the PQ
// centroid j of subquantizers [h*M/2, (h+1)*M/2) and their f64 norms in
// registers for the whole launch, so a table costs no memory traffic beyond
// the coarse centroid row.  The m x 256 FP32 table lives in LDS; every
// lane scans one vector at a time with one 8/16-byte coalesced code load and
// M LDS gathers.  Candidates that beat the current k-th key are appended to an
// LDS buffer with one LDS atomic per wavefront; when the buffer fills, an exact
// block radix select shrinks it back to k and tightens the threshold.
#include "select_dev.h"

namespace mivf {

namespace {

constexpr int D = 128;
constexpr int NT = 512;

template <int M>
struct CodeVec;
template <>
struct CodeVec<4> {
    static constexpr int W = 1;
    __device__ static void load(const uint8_t *p, uint32_t *w) { w[0] = *reinterpret_cast<const uint32_t *>(p); }
};
template <>
struct CodeVec<8> {
    static constexpr int W = 2;
    __device__ static void load(const uint8_t *p, uint32_t *w) {
        uint2 v = *reinterpret_cast<const uint2 *>(p);
        w[0] = v.x; w[1] = v.y;
    }
};
template <>
struct CodeVec<16> {
    static constexpr int W = 4;
    __device__ static void load(const uint8_t *p, uint32_t *w) {
        uint4 v = *reinterpret_cast<const uint4 *>(p);
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    }
};
template <>
struct CodeVec<32> {
    static constexpr int W = 8;
    __device__ static void load(const uint8_t *p, uint32_t *w) {
        uint4 v0 = *reinterpret_cast<const uint4 *>(p);
        uint4 v1 = *reinterpret_cast<const uint4 *>(p + 16);
        w[0] = v0.x; w[1] = v0.y; w[2] = v0.z; w[3] = v0.w;
        w[4] = v1.x; w[5] = v1.y; w[6] = v1.z; w[7] = v1.w;
    }
};

template <int CAP>
struct ScanShared {
    uint64_t cand[CAP];
    SelectShared sel;
    int cnt;
    int newcnt;
    int eqt;
};

// Shrink cand[0..cnt) to the kk smallest keys; returns the new threshold.
template <int CAP>
__device__ uint64_t compact(ScanShared<CAP> &S, int kk) {
    constexpr int PER = CAP / NT;
    const int tid = threadIdx.x;
    const int c = S.cnt;
    uint64_t *cand = S.cand;
    SelectResult r = block_select<NT>([&](int i) { return cand[i]; }, c, kk, S.sel);
    uint64_t mine[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        int i = tid + e * NT;
        mine[e] = i < c ? cand[i] : kNoKey;
    }
    if (tid == 0) {
        S.newcnt = 0;
        S.eqt = 0;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        uint64_t key = mine[e];
        if (key == kNoKey) continue;
        bool take = key < r.thr;
        if (!take && r.eq > 0 && key == r.thr) take = atomicAdd(&S.eqt, 1) < r.eq;
        if (take) {
            const int pos = atomicAdd(&S.newcnt, 1);
            if (pos < CAP) cand[pos] = key;
        }
    }
    __syncthreads();
    if (tid == 0) S.cnt = min(S.newcnt, CAP);
    __syncthreads();
    return r.thr;
}

template <int M, int KM>
__global__ __launch_bounds__(NT) void scan_kernel(ScanArgs a) {
    constexpr int DS = D / M;
    constexpr int HM = M / 2;
    constexpr int CW = CodeVec<M>::W;
    constexpr int U = 2;                 // vectors per thread per step
    constexpr int STEP = NT * U;
    constexpr int CAP = KM <= 128 ? 2048 : 4096;
    static_assert(CAP - STEP >= KM, "candidate buffer too small");

    __shared__ float tab[M * 256];
    __shared__ float qv[D];
    __shared__ float resid[D];
    __shared__ float rnorm[M];
    __shared__ ScanShared<CAP> S;

    const int tid = threadIdx.x;
    const int q = blockIdx.x / a.ngroups;
    const int g = blockIdx.x - q * a.ngroups;
    const int wq = a.wq ? a.wq[q] : a.w;
    const int kq = a.kq ? a.kq[q] : a.k;
    const int p0 = g * a.ppg;
    const int p1 = min(wq, p0 + a.ppg);

    const int j = tid & 255;
    const int h = tid >> 8;
    const bool jok = j < a.ks;  // table rows past ks are never indexed by valid codes
    float preg[HM][DS];
    double pnreg[HM];
#pragma unroll
    for (int s2 = 0; s2 < HM; ++s2) {
        const int s = h * HM + s2;
        const float *src = a.pq + (int64_t(s) * a.ks + j) * DS;
#pragma unroll
        for (int t = 0; t < DS; ++t) preg[s2][t] = jok ? src[t] : 0.0f;
        pnreg[s2] = jok ? a.pq_norm[s * a.ks + j] : 0.0;
    }
    if (!a.resid_mode && tid < D) qv[tid] = a.queries[int64_t(q) * D + tid];
    if (tid == 0) S.cnt = 0;
    uint64_t thr = kNoKey;
    uint32_t thr_hi = 0xFFFFFFFFu;
    __syncthreads();

    for (int p = p0; p < p1; ++p) {
        const int l = a.probes[int64_t(q) * a.wstride + p];
        if (l < 0 || l >= a.nlist) continue;  // uniform: every lane reads the same probe
        const int64_t beg = a.list_off[l];
        const int len = int(a.list_off[l + 1] - beg);
        if (tid < D) {
            resid[tid] = a.resid_mode ? a.queries[(int64_t(q) * a.wstride + p) * D + tid]
                                      : qv[tid] - a.coarse[int64_t(l) * D + tid];
        }
        __syncthreads();
        if (tid < M) {
            float s = 0.0f;
#pragma unroll
            for (int t = 0; t < DS; ++t) {
                float pr = resid[tid * DS + t] * resid[tid * DS + t];
                s = s + pr;
            }
            rnorm[tid] = s;
        }
        __syncthreads();
#pragma unroll
        for (int s2 = 0; s2 < HM; ++s2) {
            const int s = h * HM + s2;
            float dot = 0.0f;
#pragma unroll
            for (int t = 0; t < DS; ++t) {
                float pr = resid[s * DS + t] * preg[s2][t];
                dot = dot + pr;
            }
            float c = float(pnreg[s2] + double(rnorm[s]));
            float m2 = -2.0f * dot;
            tab[s * 256 + j] = jok ? m2 + c : __builtin_huge_valf();
        }
        __syncthreads();

        const uint8_t *lc = a.codes + beg * M;
        const int32_t *lid = a.list_ids + beg;
        for (int base = 0; base < len; base += STEP) {
            uint32_t cw[U][CW];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int v = base + u * NT + tid;
                if (v < len) {
                    CodeVec<M>::load(lc + int64_t(v) * M, cw[u]);
                } else {
#pragma unroll
                    for (int w = 0; w < CW; ++w) cw[u][w] = 0;
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                float acc = 0.0f;
#pragma unroll
                for (int w = 0; w < CW; ++w) {
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const uint32_t code = (cw[u][w] >> (8 * b)) & 255u;
                        acc = acc + tab[(w * 4 + b) * 256 + code];
                    }
                }
                const int v = base + u * NT + tid;
                const uint32_t kh = fkey(acc);
                bool pass = (v < len) && kh <= thr_hi;
                uint64_t key = 0;
                if (pass) {
                    key = (uint64_t(kh) << 32) | uint32_t(lid[v]);
                    pass = key < thr;
                }
                const uint64_t mask = __ballot(pass);
                if (mask) {
                    const int lane = __lane_id();
                    const int before = __popcll(mask & ((uint64_t(1) << lane) - 1));
                    int pos = 0;
                    if (lane == 0) pos = atomicAdd(&S.cnt, __popcll(mask));
                    pos = __shfl(pos, 0);
                    if (pass && pos + before < CAP) S.cand[pos + before] = key;
                }
            }
            __syncthreads();
            const int cur = S.cnt;
            __syncthreads();  // every lane has read cnt before the next appends
            if (cur > CAP - STEP) {
                thr = compact<CAP>(S, kq);
                thr_hi = uint32_t(thr >> 32);
            }
        }
        __syncthreads();
    }

    // final selection
    if (S.cnt > kq) compact<CAP>(S, kq);
    const int c = S.cnt;
    if (a.ngroups > 1) {
        uint64_t *dst = a.partial + (int64_t(q) * a.ngroups + g) * a.kstride;
        for (int i = tid; i < a.kstride; i += NT) dst[i] = i < c ? S.cand[i] : kNoKey;
        return;
    }
    const int np = next_pow2(c > 0 ? c : 1);
    for (int i = c + tid; i < np; i += NT) S.cand[i] = kNoKey;
    __syncthreads();
    block_bitonic_sort<NT>(S.cand, np);
    int32_t *oi = a.out_ids + int64_t(q) * a.kstride;
    float *od = a.out_dis + int64_t(q) * a.kstride;
    for (int i = tid; i < a.kstride; i += NT) {
        if (i < c) {
            oi[i] = key_id(S.cand[i]);
            od[i] = key_dist(S.cand[i]);
        } else {
            oi[i] = -1;
            od[i] = __builtin_huge_valf();
        }
    }
}

template <int M>
hipError_t launch_m(const ScanArgs &a, hipStream_t s) {
    const unsigned grid = unsigned(int64_t(a.nq) * a.ngroups);
    if (a.kstride <= 128)
        scan_kernel<M, 128><<<grid, NT, 0, s>>>(a);
    else
        scan_kernel<M, 1024><<<grid, NT, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_scan(const ScanArgs &a, int m, hipStream_t s) {
    if (a.nq <= 0) return hipSuccess;
    if (a.kstride > 1024) return hipErrorInvalidValue;
    switch (m) {
        case 4: return launch_m<4>(a, s);
        case 8: return launch_m<8>(a, s);
        case 16: return launch_m<16>(a, s);
        case 32: return launch_m<32>(a, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mivf
