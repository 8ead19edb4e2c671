// ADC distances of one residual against an arbitrary run of codes, in code
// order: the per-list search of the parallel/ fork
// (parallel/ivf_pq/ivf_search.cpp:178-228), which builds the m x ks table of
// the residual (compute_cross_distances(ds, 1, ks, r_s, P_s), :206-209) and
// sums it over each code with sumidxtab2 (:451-468).
//
// One residual per launch: every workgroup builds the whole table in LDS
// (m * ks entries, one per thread in turn, the reference arithmetic with the
// residual sub-vector as the f32-norm side and the PQ centroid as the f64-norm
// side) and then walks its share of the codes, one code per thread, summing
// the m entries sequentially in subquantizer order.
#include "mivf_internal.h"

namespace mivf {

namespace {

constexpr int NT = 256;

__global__ __launch_bounds__(NT) void adc_codes_kernel(const float *__restrict__ resid, int m, int ks, int ds,
                                                       const float *__restrict__ pq,
                                                       const double *__restrict__ pq_n64,
                                                       const uint8_t *__restrict__ codes, int64_t n,
                                                       float *__restrict__ out) {
    extern __shared__ float tab[];  // [m][ks]
    __shared__ float rs[512];
    const int tid = threadIdx.x;
    const int d = m * ds;
    for (int t = tid; t < d; t += NT) rs[t] = resid[t];
    __syncthreads();
    for (int e = tid; e < m * ks; e += NT) {
        const int s = e / ks;
        const float *a = rs + s * ds;
        const float *b = pq + int64_t(e) * ds;
        float na = 0.0f, dot = 0.0f;
        for (int t = 0; t < ds; ++t) {
            const float p = a[t] * a[t];
            na = na + p;
        }
        for (int t = 0; t < ds; ++t) {
            const float p = a[t] * b[t];
            dot = dot + p;
        }
        const float c = float(pq_n64[e] + double(na));
        const float m2 = -2.0f * dot;
        tab[e] = m2 + c;
    }
    __syncthreads();
    for (int64_t i = int64_t(blockIdx.x) * NT + tid; i < n; i += int64_t(gridDim.x) * NT) {
        const uint8_t *cv = codes + i * m;
        float sum = 0.0f;
        for (int j = 0; j < m; ++j) sum = sum + tab[j * ks + cv[j]];
        out[i] = sum;
    }
}

}  // namespace

hipError_t launch_adc_codes(const float *resid, int m, int ks, int ds, const float *pq, const double *pq_n64,
                            const uint8_t *codes, int64_t n, float *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (m < 1 || ks < 1 || ks > 256 || ds < 1 || m * ds > 512 || m * ks > 8192) return hipErrorInvalidValue;
    const int64_t blocks = std::min<int64_t>((n + NT - 1) / NT, 1024);
    adc_codes_kernel<<<unsigned(blocks), NT, sizeof(float) * size_t(m) * ks, s>>>(resid, m, ks, ds, pq, pq_n64,
                                                                                 codes, n, out);
    return hipGetLastError();
}

}  // namespace mivf
