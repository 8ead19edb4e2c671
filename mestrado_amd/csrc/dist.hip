// Exact L2 distances in the reference's arithmetic (parallel/yael_needs/nn.c:121-150
// with the netlib SGEMM loop order): for a-row i and b-row j
//     c   = (float)(nb_j + (double)na_i)      na: f32 accumulator, nb: f64 accumulator
//     dot = sum_t fl32(a_it * b_jt)           t ascending, one rounding per step, no FMA
//     D   = fl32(fl32(-2 * dot) + c)
// The tile kernel is a register-blocked product whose per-output accumulation
// runs over t in order, so each output is the same sequential chain the CPU
// computes.  Used for coarse assignment (a = coarse centroids, b = queries),
// k-means assignment and PQ encoding (argmin epilogue).
#include "mivf_internal.h"

namespace mivf {

namespace {

constexpr int TA = 64;   // a rows per tile
constexpr int TB = 64;   // b rows per tile
constexpr int KC = 16;   // k chunk staged in LDS
constexpr int NT = 256;  // threads: 16 x 16, each owning 4 x 4 outputs

__global__ void norm_f32_kernel(const float *__restrict__ x, int64_t n, int d, int ldx,
                                float *__restrict__ out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = x + i * ldx;
    float s = 0.0f;
    for (int t = 0; t < d; ++t) {
        float p = r[t] * r[t];
        s = s + p;
    }
    out[i] = s;
}

__global__ void norm_f64_kernel(const float *__restrict__ x, int64_t n, int d, int ldx,
                                double *__restrict__ out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = x + i * ldx;
    double s = 0.0;
    for (int t = 0; t < d; ++t) {
        float p = r[t] * r[t];
        s = s + double(p);
    }
    out[i] = s;
}

template <bool ARGMIN>
__global__ __launch_bounds__(NT) void dist_tile_kernel(
    const float *__restrict__ a, int lda, const float *__restrict__ an, int na,
    const float *__restrict__ b, int ldb, const double *__restrict__ bn, int64_t nb,
    int d, float *__restrict__ out, int64_t ldo, uint64_t *__restrict__ keys) {
    __shared__ float As[KC][TA];
    __shared__ float Bs[KC][TB];

    const int tid = threadIdx.x;
    const int tx = tid & 15;   // a direction
    const int ty = tid >> 4;   // b direction
    const int a0 = blockIdx.x * TA;
    const int64_t b0 = int64_t(blockIdx.y) * TB;

    float acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0f;

    for (int k0 = 0; k0 < d; k0 += KC) {
        const int kc = min(KC, d - k0);
        // stage a tile (TA x KC) and b tile (TB x KC) transposed to k-major
#pragma unroll
        for (int e4 = 0; e4 < (TA * KC) / NT; ++e4) {
            int e = tid + e4 * NT;
            int row = e / KC, col = e % KC;
            float va = 0.0f, vb = 0.0f;
            if (col < kc) {
                if (a0 + row < na) va = a[int64_t(a0 + row) * lda + k0 + col];
                if (b0 + row < nb) vb = b[(b0 + row) * ldb + k0 + col];
            }
            As[col][row] = va;
            Bs[col][row] = vb;
        }
        __syncthreads();
        for (int kk = 0; kk < kc; ++kk) {
            const float4 av = *reinterpret_cast<const float4 *>(&As[kk][tx * 4]);
            const float4 bv = *reinterpret_cast<const float4 *>(&Bs[kk][ty * 4]);
            const float ar[4] = {av.x, av.y, av.z, av.w};
            const float br[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    float p = ar[v] * br[u];
                    acc[u][v] = acc[u][v] + p;
                }
        }
        __syncthreads();
    }

#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t j = b0 + ty * 4 + u;
        const bool jok = j < nb;
        const double bnj = jok ? bn[j] : 0.0;
        uint64_t best = kNoKey;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = a0 + tx * 4 + v;
            if (jok && i < na) {
                float c = float(bnj + double(an[i]));
                float m2 = -2.0f * acc[u][v];
                float D = m2 + c;
                if constexpr (ARGMIN) {
                    uint64_t kk = ckey(D, i);
                    best = kk < best ? kk : best;
                } else {
                    out[j * ldo + i] = D;
                }
            }
        }
        if constexpr (ARGMIN) {
            // reduce over the 16 lanes (tx) that share this b row
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) {
                uint64_t o = __shfl_xor(best, off, 16);
                best = o < best ? o : best;
            }
            if (tx == 0 && jok && best != kNoKey) atomicMin(reinterpret_cast<unsigned long long *>(keys + j), (unsigned long long)best);
        }
    }
}

__global__ void fill_u64_kernel(uint64_t *p, int64_t n, uint64_t v) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void keys_to_idx_kernel(const uint64_t *keys, int64_t n, int32_t *idx) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) idx[i] = keys[i] == kNoKey ? -1 : key_id(keys[i]);
}

__global__ void codes_from_keys_kernel(const uint64_t *keys, int64_t n, int m, int sub,
                                       uint8_t *codes) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) codes[i * m + sub] = uint8_t(key_id(keys[i]) & 0xFF);
}

__global__ void residual_kernel(const float *__restrict__ x, int64_t n, int d,
                                const float *__restrict__ c, const int32_t *__restrict__ assign,
                                float *__restrict__ r) {
    int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * d) return;
    int64_t i = e / d;
    int t = int(e % d);
    r[e] = x[e] - c[int64_t(assign[i]) * d + t];
}

inline unsigned nblk(int64_t n, int t) { return unsigned((n + t - 1) / t); }

}  // namespace

hipError_t launch_norm_f32(const float *x, int64_t n, int d, int ldx, float *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    norm_f32_kernel<<<nblk(n, 256), 256, 0, s>>>(x, n, d, ldx, out);
    return hipGetLastError();
}

// Same sums, rows staged through LDS: 16 rows per workgroup (a 10k-query
// batch spreads over 625 workgroups), 32 dimensions per chunk loaded
// coalesced by all 256 threads, then each of 16 threads continues its row's
// sequential f64 sum over the chunk (t ascending).
constexpr int NR = 16, NCH = 32;
__global__ __launch_bounds__(256) void norm_f64_rows_kernel(const float *__restrict__ x, int64_t n, int d, int ldx,
                                                            double *__restrict__ out) {
    __shared__ float tile[NR][NCH + 1];
    const int tid = threadIdx.x;
    const int64_t r0 = int64_t(blockIdx.x) * NR;
    double s = 0.0;
    for (int t0 = 0; t0 < d; t0 += NCH) {
        for (int e = tid; e < NR * NCH; e += 256) {
            const int rr = e / NCH, tt = e % NCH;
            const int64_t row = r0 + rr;
            tile[rr][tt] = (row < n && t0 + tt < d) ? x[row * ldx + t0 + tt] : 0.0f;
        }
        __syncthreads();
        if (tid < NR) {
            const int nt = min(NCH, d - t0);
            for (int tt = 0; tt < nt; ++tt) {
                const float v = tile[tid][tt];
                const float p = v * v;
                s = s + double(p);
            }
        }
        __syncthreads();
    }
    if (tid < NR && r0 + tid < n) out[r0 + tid] = s;
}

hipError_t launch_norm_f64(const float *x, int64_t n, int d, int ldx, double *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n < (int64_t(1) << 20)) {  // small batches (queries): coalesced through LDS
        norm_f64_rows_kernel<<<nblk(n, NR), 256, 0, s>>>(x, n, d, ldx, out);
    } else {
        norm_f64_kernel<<<nblk(n, 256), 256, 0, s>>>(x, n, d, ldx, out);
    }
    return hipGetLastError();
}

hipError_t launch_dist_store(const float *a, int lda, const float *an, int na,
                             const float *b, int ldb, const double *bn, int64_t nb,
                             int d, float *out, int64_t ldo, hipStream_t s) {
    if (na <= 0 || nb <= 0) return hipSuccess;
    dim3 grid(nblk(na, TA), nblk(nb, TB));
    dist_tile_kernel<false><<<grid, NT, 0, s>>>(a, lda, an, na, b, ldb, bn, nb, d, out, ldo,
                                                nullptr);
    return hipGetLastError();
}

hipError_t launch_dist_argmin(const float *a, int lda, const float *an, int na,
                              const float *b, int ldb, const double *bn, int64_t nb,
                              int d, uint64_t *keys, hipStream_t s) {
    if (na <= 0 || nb <= 0) return hipSuccess;
    dim3 grid(nblk(na, TA), nblk(nb, TB));
    dist_tile_kernel<true><<<grid, NT, 0, s>>>(a, lda, an, na, b, ldb, bn, nb, d, nullptr, 0,
                                               keys);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t *p, int64_t n, uint64_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    fill_u64_kernel<<<nblk(n, 256), 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_keys_to_idx(const uint64_t *keys, int64_t n, int32_t *idx, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    keys_to_idx_kernel<<<nblk(n, 256), 256, 0, s>>>(keys, n, idx);
    return hipGetLastError();
}

hipError_t launch_codes_from_keys(const uint64_t *keys, int64_t n, int m, int sub,
                                  uint8_t *codes, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    codes_from_keys_kernel<<<nblk(n, 256), 256, 0, s>>>(keys, n, m, sub, codes);
    return hipGetLastError();
}

hipError_t launch_residual(const float *x, int64_t n, int d, const float *c,
                           const int32_t *assign, float *r, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    residual_kernel<<<nblk(n * d, 256), 256, 0, s>>>(x, n, d, c, assign, r);
    return hipGetLastError();
}

}  // namespace mivf
