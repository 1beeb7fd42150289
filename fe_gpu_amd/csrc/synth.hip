// Device generator of the worker-firm-time benchmark panel: n_workers x n_periods rows ordered by
// worker then period; each worker starts at a uniform firm and moves to a uniform firm with
// probability switch_p per period; tenure counts periods since joining the current firm;
// sick_shock ~ Bernoulli(0.05); extra_1..6 = 0.3 a_w + sqrt(0.91) e with e ~ N(0,1);
// y = X beta + a_w + a_f + a_t + eps with a_w ~ N(0,1), a_f ~ N(0,0.5), a_t ~ N(0,0.1), eps ~ N(0,1).
// Random numbers come from a counter-based hash of (seed, stream, index), so every row is a
// pure function of its coordinates and the panel is identical for any launch shape.
#include "../../include/fehip.h"
#include "common.hpp"

namespace fehip {
namespace {

__device__ __forceinline__ uint64_t mix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double uni(uint64_t seed, uint64_t stream, uint64_t idx) {
    const uint64_t h = mix(mix(seed ^ (stream * 0xd1b54a32d192ed03ULL)) ^ idx);
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);  // (0,1)
}

__device__ __forceinline__ double gauss(uint64_t seed, uint64_t stream, uint64_t idx) {
    const double u1 = uni(seed, stream, 2 * idx), u2 = uni(seed, stream, 2 * idx + 1);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__constant__ double kBeta[8] = {0.4, -0.2, -0.2665779136218663, 0.06381245236427056, 0.4631683295023366,
                                0.7596569511243322, 0.9276463627630034, -0.09644476806482337};

__global__ void synth_kernel(int64_t n_workers, int T, int n_firms, double switch_p, uint64_t seed, int32_t* fe,
                             double* y, double* W) {
    const int64_t n = n_workers * T;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n_workers;
         w += (int64_t)gridDim.x * blockDim.x) {
        const double aw = gauss(seed, 1, (uint64_t)w);
        int firm = 1 + (int)(uni(seed, 2, (uint64_t)w * 64) * n_firms);
        if (firm > n_firms) firm = n_firms;
        double tenure = 0.0;
        for (int t = 0; t < T; ++t) {
            const int64_t i = w * T + t;
            if (t > 0) {
                if (uni(seed, 3, (uint64_t)i) < switch_p) {
                    int nf = 1 + (int)(uni(seed, 4, (uint64_t)i) * n_firms);
                    firm = nf > n_firms ? n_firms : nf;
                    tenure = 0.0;
                } else {
                    tenure += 1.0;
                }
            }
            const double af = 0.5 * gauss(seed, 5, (uint64_t)firm);
            const double at = 0.1 * gauss(seed, 6, (uint64_t)t);
            const double sick = uni(seed, 7, (uint64_t)i) < 0.05 ? 1.0 : 0.0;
            double xb = kBeta[0] * tenure + kBeta[1] * sick;
            W[i] = tenure;
            W[i + n] = sick;
            for (int j = 0; j < 6; ++j) {
                const double ex = 0.3 * aw + 0.9539392014169456 * gauss(seed, 8 + j, (uint64_t)i);
                W[i + (int64_t)(2 + j) * n] = ex;
                xb += kBeta[2 + j] * ex;
            }
            y[i] = xb + aw + af + at + gauss(seed, 20, (uint64_t)i);
            fe[i] = (int32_t)(w + 1);
            fe[i + n] = firm;
            fe[i + 2 * n] = t + 1;
        }
    }
}

}  // namespace
}  // namespace fehip

using namespace fehip;

extern "C" int fehip_synth_worker_firm(int64_t n_workers, int32_t n_periods, int32_t n_firms, double switch_p,
                                       uint64_t seed, int32_t* d_fe, double* d_y, double* d_W) {
    if (n_workers <= 0 || n_periods <= 0 || n_firms <= 0 || !d_fe || !d_y || !d_W)
        return set_error("fehip_synth_worker_firm: bad arguments");
    if (n_workers * (int64_t)n_periods > 0x7fffffffLL) return set_error("fehip_synth_worker_firm: too many rows");
    hipLaunchKernelGGL(synth_kernel, dim3(stream_grid((size_t)n_workers, 256)), dim3(256), 0, 0, n_workers, n_periods,
                       n_firms, switch_p, seed, d_fe, d_y, d_W);
    FEHIP_LAUNCH_CHECK("synth kernel");
    FEHIP_CHECK(hipDeviceSynchronize(), "synth sync");
    return set_ok();
}

// Plain streaming copy (16-byte vectors), the reference kernel for calibrating HBM byte counters.
__global__ void device_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int fehip_device_copy(const void* d_src, void* d_dst, int64_t bytes) {
    (void)hipGetLastError();
    if (!d_src || !d_dst || bytes < 0 || (bytes & 15)) return fehip::set_error("fehip_device_copy: bad arguments");
    const int64_t n16 = bytes / 16;
    if (n16 == 0) return fehip::set_ok();
    hipLaunchKernelGGL(device_copy_kernel, dim3(fehip::stream_grid((size_t)n16, 256)), dim3(256), 0, 0,
                       (const uint4*)d_src, (uint4*)d_dst, n16);
    FEHIP_LAUNCH_CHECK("device copy");
    FEHIP_CHECK(hipDeviceSynchronize(), "device copy");
    return fehip::set_ok();
}
