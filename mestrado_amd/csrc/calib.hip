// LDS read-rate calibration: the ceiling that binds the table-gather loops
// of the scan kernels.  Every lane reads `bytes` (4 or 16) per step from a
// 64 KiB LDS table, for the same 16 subquantizer blocks of 256 rows the scan
// kernels use, with either random rows (a per-lane LCG: the scan's data-driven
// gathers, bank conflicts included) or rows that keep every wave's reads on
// distinct banks (the conflict-free peak).  Workgroups of 512 threads, two per
// CU, as the list-major exact scan runs.
#include "mivf_internal.h"

namespace mivf {
namespace {

constexpr int NT = 512;
constexpr int STEPS = 4096;
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int BYTES, bool RANDOM>
__global__ __launch_bounds__(NT, 4) void lds_rate_kernel(float *__restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    float *t = reinterpret_cast<float *>(smem);
    for (int i = tid; i < 16384; i += NT) t[i] = float(i & 1023);
    __syncthreads();
    uint32_t x = 2654435761u * uint32_t(tid + 1) + uint32_t(blockIdx.x) * 40503u;
    float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, acc3 = 0.0f;
    const int lane = tid & 63;
    for (int i = 0; i < STEPS; ++i) {
        uint32_t row;
        if (RANDOM) {
            x = x * 1664525u + 1013904223u;
            row = x >> 24;  // 0..255
        } else {
            row = uint32_t(lane + 64 * (i & 3));  // consecutive rows: distinct banks
        }
        const uint32_t blk = uint32_t(i & 15) * 256u;
        if (BYTES == 16) {
            const f32x4 v = *reinterpret_cast<__attribute__((address_space(3))) const f32x4 *>(
                uintptr_t((blk + row) * 16u));
            acc0 = acc0 + v[0];
            acc1 = acc1 + v[1];
            acc2 = acc2 + v[2];
            acc3 = acc3 + v[3];
        } else {
            const float v = *reinterpret_cast<__attribute__((address_space(3))) const float *>(
                uintptr_t((blk + row) * 4u));
            acc0 = acc0 + v;
        }
    }
    const float s = acc0 + acc1 + acc2 + acc3;
    if (s == -1.0f) out[blockIdx.x * NT + tid] = s;  // keeps the reads; never true
}

template <int BYTES, bool RANDOM>
int time_rate(int ncu, double *gbps) {
    const size_t lds = 65536;
    MIVF_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&lds_rate_kernel<BYTES, RANDOM>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    const int grid = 2 * ncu;
    float *out = nullptr;
    MIVF_CHECK_HIP(hipMalloc(&out, sizeof(float) * size_t(grid) * NT));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = kOk;
    float best = 1e30f;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = kHip;
    for (int rep = 0; rep < 4 && rc == kOk; ++rep) {
        lds_rate_kernel<BYTES, RANDOM><<<grid, NT, lds, nullptr>>>(out);
        if (hipGetLastError() != hipSuccess) { rc = kHip; break; }
        if (hipEventRecord(e0, nullptr) != hipSuccess) { rc = kHip; break; }
        lds_rate_kernel<BYTES, RANDOM><<<grid, NT, lds, nullptr>>>(out);
        if (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = kHip; break; }
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = kHip; break; }
        best = std::min(best, ms);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (rc != kOk) {
        set_error("lds calibration: HIP event or launch failure");
        return rc;
    }
    *gbps = double(grid) * NT * STEPS * BYTES / (double(best) * 1e-3) / 1e9;
    return kOk;
}

}  // namespace
}  // namespace mivf

using namespace mivf;

extern "C" int mivf_lds_read_rate(int bytes, int random_rows, double *gbps) try {
    MIVF_REQUIRE(gbps && (bytes == 4 || bytes == 16), kInvalid, "lds_read_rate: bytes must be 4 or 16");
    int dev = 0, ncu = 0;
    MIVF_CHECK_HIP(hipGetDevice(&dev));
    MIVF_CHECK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (bytes == 16) return random_rows ? time_rate<16, true>(ncu, gbps) : time_rate<16, false>(ncu, gbps);
    return random_rows ? time_rate<4, true>(ncu, gbps) : time_rate<4, false>(ncu, gbps);
} MIVF_CATCH_ALL
