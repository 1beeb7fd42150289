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

// Wide 2SLS design (20M-row shape): individuals x periods rows ordered by individual then period,
// firms moving with probability switch_p per period, time = period.  One thread per row (the firm
// path of its individual is replayed up to the row, so every row is a pure function of its
// coordinates and all column writes are coalesced).  Columns of W (K = 252):
//   0 xend (endogenous), 1-7 cat8 levels 2..8 (endogenous), 8-11 x1..x4,
//   12-40 ind levels 2..30, 41-47 msp levels 2..8, 48-250 ind x msp (levels 2.. x 2..),
//   251 x1 * x2;
// Z (L = 21): 0 z1, 1 z2, 2-20 zcat20 levels 2..20.  xend and the cat8 latent load on the
// instruments (first-stage correlation about 0.4) and on the structural error u.
constexpr int kWideK = 252;
constexpr int kWideL = 21;

__global__ void synth_iv_wide_kernel(int64_t n_ind, int T, int n_firms, double switch_p, uint64_t seed, int32_t* fe,
                                     double* y, double* W, double* Z) {
    const int64_t n = n_ind * T;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t w = i / T;
        const int t = (int)(i - w * T);
        int firm = 1 + (int)(uni(seed, 2, (uint64_t)w * 64) * n_firms);
        for (int s = 1; s <= t; ++s)
            if (uni(seed, 3, (uint64_t)(w * T + s)) < switch_p) firm = 1 + (int)(uni(seed, 4, (uint64_t)(w * T + s)) * n_firms);
        if (firm > n_firms) firm = n_firms;
        const double z1 = gauss(seed, 30, (uint64_t)i), z2 = gauss(seed, 31, (uint64_t)i);
        int zc = 1 + (int)(uni(seed, 32, (uint64_t)i) * 20);
        zc = zc > 20 ? 20 : zc;
        const double u = gauss(seed, 33, (uint64_t)i);
        const double xend = 0.45 * 0.7071067811865476 * (z1 + z2) + 0.6 * u + 0.66 * gauss(seed, 34, (uint64_t)i);
        const double lat = 0.4 * ((double)zc - 10.5) / 5.766 + 0.5 * u + 0.77 * gauss(seed, 35, (uint64_t)i);
        int c8 = 1 + (int)floor((lat + 2.5) * 1.6);
        c8 = c8 < 1 ? 1 : (c8 > 8 ? 8 : c8);
        double x[4];
        for (int j = 0; j < 4; ++j) x[j] = gauss(seed, 36 + j, (uint64_t)i);
        int ind = 1 + (int)(uni(seed, 40, (uint64_t)i) * 30);
        ind = ind > 30 ? 30 : ind;
        int msp = 1 + (int)(uni(seed, 41, (uint64_t)i) * 8);
        msp = msp > 8 ? 8 : msp;
        // columns
        for (int c = 0; c < kWideK; ++c) W[i + (int64_t)c * n] = 0.0;
        W[i] = xend;
        if (c8 >= 2) W[i + (int64_t)(c8 - 1) * n] = 1.0;
        for (int j = 0; j < 4; ++j) W[i + (int64_t)(8 + j) * n] = x[j];
        if (ind >= 2) W[i + (int64_t)(12 + ind - 2) * n] = 1.0;
        if (msp >= 2) W[i + (int64_t)(41 + msp - 2) * n] = 1.0;
        if (ind >= 2 && msp >= 2) W[i + (int64_t)(48 + (ind - 2) * 7 + (msp - 2)) * n] = 1.0;
        W[i + (int64_t)251 * n] = x[0] * x[1];
        for (int c = 0; c < kWideL; ++c) Z[i + (int64_t)c * n] = 0.0;
        Z[i] = z1;
        Z[i + n] = z2;
        if (zc >= 2) Z[i + (int64_t)(2 + zc - 2) * n] = 1.0;
        // outcome: effects of xend, the cat8 levels, x1..x4, ind, msp, their interaction and x1 x2
        const double aw = gauss(seed, 1, (uint64_t)w), af = 0.5 * gauss(seed, 5, (uint64_t)firm);
        const double at = 0.1 * gauss(seed, 6, (uint64_t)t);
        double v = 0.8 * xend + 0.15 * (double)(c8 - 1) + 0.5 * x[0] - 0.3 * x[1] + 0.2 * x[2] + 0.1 * x[3];
        v += 0.02 * (double)ind - 0.05 * (double)msp + 0.01 * (double)((ind * msp) % 7) + 0.25 * x[0] * x[1];
        y[i] = v + aw + af + at + u + gauss(seed, 42, (uint64_t)i);
        fe[i] = (int32_t)(w + 1);
        fe[i + n] = firm;
        fe[i + 2 * n] = t + 1;
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

extern "C" int fehip_synth_iv_wide(int64_t n_ind, int32_t n_periods, int32_t n_firms, double switch_p, uint64_t seed,
                                   int32_t* d_fe, double* d_y, double* d_W, double* d_Z) {
    (void)hipGetLastError();
    if (n_ind <= 0 || n_periods <= 0 || n_firms <= 0 || !d_fe || !d_y || !d_W || !d_Z)
        return fehip::set_error("fehip_synth_iv_wide: bad arguments");
    const int64_t n = n_ind * n_periods;
    if (n > 0x7fffffffLL) return fehip::set_error("fehip_synth_iv_wide: too many rows");
    hipLaunchKernelGGL(fehip::synth_iv_wide_kernel, dim3(fehip::stream_grid((size_t)n, 256)), dim3(256), 0, 0, n_ind,
                       (int)n_periods, (int)n_firms, switch_p, seed, d_fe, d_y, d_W, d_Z);
    FEHIP_LAUNCH_CHECK("iv wide generator");
    FEHIP_CHECK(hipDeviceSynchronize(), "iv wide generator");
    return fehip::set_ok();
}

namespace fehip {
namespace {
__global__ void synth_mask_kernel(int64_t n, double p, uint64_t seed, uint8_t* m) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m[i] = uni(seed, 50, (uint64_t)i) < p ? 1 : 0;
}
}  // namespace
}  // namespace fehip

extern "C" int fehip_synth_mask(int64_t n, double p, uint64_t seed, uint8_t* d_mask) {
    (void)hipGetLastError();
    if (n < 0 || (n > 0 && !d_mask) || !(p >= 0.0 && p <= 1.0)) return fehip::set_error("fehip_synth_mask: bad arguments");
    if (n == 0) return fehip::set_ok();
    hipLaunchKernelGGL(fehip::synth_mask_kernel, dim3(fehip::stream_grid((size_t)n, 256)), dim3(256), 0, 0, n, p, seed,
                       d_mask);
    FEHIP_LAUNCH_CHECK("mask generator");
    FEHIP_CHECK(hipStreamSynchronize(0), "mask generator");  // complete before other threads' streams read it
    return fehip::set_ok();
}
