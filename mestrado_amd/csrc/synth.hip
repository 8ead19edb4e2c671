// Synthetic data and exact ground truth on the device.
//
// gen_rows: the seeded Gaussian-mixture generator used for every workload
// (d-dim integer-valued float32 rows in [0, 255]).  SplitMix64 is used as a
// counter-based hash; a normal deviate is the Irwin-Hall sum of twelve 16-bit
// uniforms, so every step is integer arithmetic or a correctly rounded IEEE
// operation and the rows are bit-identical to the host generator.
//
// bruteforce_nn: exact FP64 nearest neighbour of each query (ties -> lowest
// id), the ground truth behind recall@R (parallel/pq-utils/pq_test_compute_stats.cpp:13-72).
#include "mivf_internal.h"

namespace mivf {

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}
__host__ __device__ inline uint64_t draw(uint64_t base, uint64_t ctr) {
    return mix64(base + (ctr + 1) * kGolden);
}

__global__ void gen_rows_kernel(uint64_t base, int64_t row0, int64_t n, int d,
                                const float *__restrict__ centres, int ncomp,
                                const uint64_t *__restrict__ thr, float sigma,
                                float *__restrict__ out) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * d) return;
    const int64_t rr = e / d;
    const int i = int(e - rr * d);
    const uint64_t r = uint64_t(row0 + rr);
    const uint64_t per_row = 1 + 3 * uint64_t(d);
    const uint64_t u53 = draw(base, r * per_row) >> 11;
    int comp;
    if (!thr) {
        comp = int((u53 * uint64_t(ncomp)) >> 53);
    } else {
        // smallest c with u53 < thr[c] (thr is non-decreasing, thr[ncomp-1] = 2^53)
        int lo = 0, hi = ncomp - 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (u53 < thr[mid]) hi = mid;
            else lo = mid + 1;
        }
        comp = lo;
    }
    int64_t s = 0;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        uint64_t g = draw(base, r * per_row + 1 + 3 * uint64_t(i) + t);
        s += int64_t(g & 0xFFFF) + int64_t((g >> 16) & 0xFFFF) + int64_t((g >> 32) & 0xFFFF) +
             int64_t(g >> 48);
    }
    const double z = double(s - 393210) / 65536.0;
    const double x = double(centres[int64_t(comp) * d + i]) + double(sigma) * z;
    double v = __builtin_rint(x);
    v = v < 0.0 ? 0.0 : v;
    v = v > 255.0 ? 255.0 : v;
    out[e] = float(v);
}

// Thread = query (query row held in registers), base rows staged through LDS
// in tiles of 32 and read as broadcasts.  Grid = (query blocks, base chunks);
// each block writes its chunk's best (distance, id) per query.
constexpr int BF_NT = 256;
constexpr int BF_TILE = 32;

template <int DD>
__global__ __launch_bounds__(BF_NT) void bruteforce_kernel(const float *__restrict__ base,
                                                           int64_t nb, const float *__restrict__ q,
                                                           int nq, int64_t chunk,
                                                           double *__restrict__ part_d,
                                                           int32_t *__restrict__ part_i) {
    __shared__ float tile[BF_TILE][DD];
    const int tid = threadIdx.x;
    const int qi = blockIdx.x * BF_NT + tid;
    const int64_t c0 = int64_t(blockIdx.y) * chunk;
    const int64_t c1 = c0 + chunk < nb ? c0 + chunk : nb;
    float qr[DD];
#pragma unroll
    for (int t = 0; t < DD; ++t) qr[t] = qi < nq ? q[int64_t(qi) * DD + t] : 0.0f;
    double best = __builtin_huge_val();
    int32_t bid = -1;
    for (int64_t t0 = c0; t0 < c1; t0 += BF_TILE) {
        const int nt = int(c1 - t0 < BF_TILE ? c1 - t0 : BF_TILE);
        __syncthreads();
        for (int e = tid; e < BF_TILE * DD; e += BF_NT) {
            int r = e / DD, c = e % DD;
            tile[r][c] = r < nt ? base[(t0 + r) * DD + c] : 0.0f;
        }
        __syncthreads();
        for (int r = 0; r < nt; ++r) {
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < DD; ++t) {
                double df = double(qr[t]) - double(tile[r][t]);
                double sq = df * df;
                s = s + sq;
            }
            if (s < best) {
                best = s;
                bid = int32_t(t0 + r);
            }
        }
    }
    if (qi < nq) {
        part_d[int64_t(blockIdx.y) * nq + qi] = best;
        part_i[int64_t(blockIdx.y) * nq + qi] = bid;
    }
}

__global__ void bruteforce_reduce_kernel(const double *part_d, const int32_t *part_i, int nq,
                                         int nchunks, int32_t *ids, double *dis) {
    const int qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    double best = __builtin_huge_val();
    int32_t bid = -1;
    for (int c = 0; c < nchunks; ++c) {  // chunks in id order: strict '<' keeps the lowest id
        double v = part_d[int64_t(c) * nq + qi];
        if (v < best) {
            best = v;
            bid = part_i[int64_t(c) * nq + qi];
        }
    }
    ids[qi] = bid;
    dis[qi] = best;
}

}  // namespace

uint64_t synth_stream_base(uint64_t seed, uint64_t stream) {
    return mix64(seed ^ mix64(stream + kGolden));
}
uint64_t synth_draw(uint64_t base, uint64_t ctr) { return draw(base, ctr); }

hipError_t launch_gen_rows(uint64_t seed, uint64_t stream, int64_t row0, int64_t n, int d,
                           const float *centres, int ncomp, const uint64_t *thr,
                           float sigma, float *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t tot = n * d;
    gen_rows_kernel<<<unsigned((tot + 255) / 256), 256, 0, s>>>(
        synth_stream_base(seed, stream), row0, n, d, centres, ncomp, thr, sigma, out);
    return hipGetLastError();
}

hipError_t launch_bruteforce_nn(const float *base, int64_t nb, const float *q, int nq,
                                int d, int32_t *ids, double *dis, double *part_d,
                                int32_t *part_i, int nchunks, hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    if (d != 128) return hipErrorInvalidValue;
    const int64_t chunk = (nb + nchunks - 1) / nchunks;
    dim3 grid(unsigned((nq + BF_NT - 1) / BF_NT), unsigned(nchunks));
    bruteforce_kernel<128><<<grid, BF_NT, 0, s>>>(base, nb, q, nq, chunk, part_d, part_i);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    bruteforce_reduce_kernel<<<unsigned((nq + 255) / 256), 256, 0, s>>>(part_d, part_i, nq,
                                                                        nchunks, ids, dis);
    return hipGetLastError();
}

}  // namespace mivf
