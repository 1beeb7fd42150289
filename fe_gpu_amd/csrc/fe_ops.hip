// FE projection pieces of the drop-in C-ABI: group accumulation, means, relaxation mix,
// subtraction, |max| and column gathers.  Behaviour follows the reference entry points
// (src/gpu/fe_gpu_runtime.cu:573-698); the kernels are written for CDNA4:
//   * accumulate privatises small group tables in LDS (one flush per workgroup) and, for
//     large tables, merges runs of equal ids inside each 64-lane wavefront before the
//     global FP64 atomic, so sorted or sticky ids cost one atomic per run, not per row;
//   * the reductions use wavefront shuffles and a single device-side atomic max.
#include <cstring>

#include "../../include/fe_gpu_runtime.h"
#include "common.hpp"

namespace fehip {
namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------------------
// accumulate, small G: LDS tables [col][g] + counts, flushed once per workgroup.
__global__ __launch_bounds__(kBlock) void accumulate_lds_kernel(
    const double* __restrict__ y, const double* __restrict__ W, const double* __restrict__ Z,
    const int* __restrict__ ids, size_t n, int G, int K, int L, size_t ld, double* __restrict__ sy,
    double* __restrict__ sW, double* __restrict__ sZ, int* __restrict__ cnt) {
    extern __shared__ double lds[];
    const int C = 1 + K + L;
    double* tab = lds;                          // C * G doubles, [c * G + g]
    int* lcnt = reinterpret_cast<int*>(lds + (size_t)C * G);
    for (int e = threadIdx.x; e < C * G; e += blockDim.x) tab[e] = 0.0;
    for (int e = threadIdx.x; e < G; e += blockDim.x) lcnt[e] = 0;
    __syncthreads();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int g = ids[i] - 1;
        if (g < 0 || g >= G) continue;  // ids outside 1..n_groups are skipped, as the reference's small-G kernel does
        atomicAdd(&tab[g], y[i]);
        for (int k = 0; k < K; ++k) atomicAdd(&tab[(1 + k) * G + g], W[i + (size_t)k * ld]);
        for (int k = 0; k < L; ++k) atomicAdd(&tab[(1 + K + k) * G + g], Z[i + (size_t)k * ld]);
        atomicAdd(&lcnt[g], 1);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < C * G; e += blockDim.x) {
        double v = tab[e];
        if (v == 0.0) continue;
        int c = e / G, g = e - c * G;
        double* dst = (c == 0) ? &sy[g] : (c <= K ? &sW[(size_t)g * K + (c - 1)] : &sZ[(size_t)g * L + (c - 1 - K)]);
        unsafeAtomicAdd(dst, v);
    }
    for (int g = threadIdx.x; g < G; g += blockDim.x)
        if (lcnt[g]) atomicAdd(&cnt[g], lcnt[g]);
}

// Sum of v over the lanes [lane, seg_end] of the calling lane's run (shuffle-down segmented
// reduction: after step k a lane holds the sum of up to 2^k following lanes of its run).
__device__ __forceinline__ double run_sum(double v, int lane, int seg_end) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        double t = __shfl_down(v, off, kWave);
        if (lane + off <= seg_end) v += t;
    }
    return v;
}

// accumulate, large G: one row per lane, runs of equal ids merged within the wavefront.
__global__ __launch_bounds__(kBlock) void accumulate_runs_kernel(
    const double* __restrict__ y, const double* __restrict__ W, const double* __restrict__ Z,
    const int* __restrict__ ids, size_t n, int G, int K, int L, size_t ld, double* __restrict__ sy,
    double* __restrict__ sW, double* __restrict__ sZ, int* __restrict__ cnt) {
    const int lane = threadIdx.x & (kWave - 1);
    const size_t wave_stride = (size_t)gridDim.x * blockDim.x;
    for (size_t base = (size_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(kWave - 1)); base < n;
         base += wave_stride) {
        const size_t i = base + lane;
        const bool in = i < n;
        int g = in ? ids[i] - 1 : -1;
        // ids outside 1..n_groups form singleton runs and are skipped (the reference's large-G kernel
        // would write past its tables for ids above n_groups; its small-G kernel skips them)
        if (g < 0 || g >= G) g = -1 - lane;
        int gnext = __shfl_down(g, 1, kWave);
        int gprev = __shfl_up(g, 1, kWave);
        bool tail = (lane == kWave - 1) || (gnext != g);
        bool head = (lane == 0) || (gprev != g);
        unsigned long long tails = __ballot(tail);
        unsigned long long above = tails & (~0ull << lane);
        int seg_end = __ffsll((long long)above) - 1;
        bool emit = head && g >= 0;
        double v = run_sum(in && g >= 0 ? y[i] : 0.0, lane, seg_end);
        if (emit) unsafeAtomicAdd(&sy[g], v);
        for (int k = 0; k < K; ++k) {
            v = run_sum(in && g >= 0 ? W[i + (size_t)k * ld] : 0.0, lane, seg_end);
            if (emit) unsafeAtomicAdd(&sW[(size_t)g * K + k], v);
        }
        for (int k = 0; k < L; ++k) {
            v = run_sum(in && g >= 0 ? Z[i + (size_t)k * ld] : 0.0, lane, seg_end);
            if (emit) unsafeAtomicAdd(&sZ[(size_t)g * L + k], v);
        }
        if (emit) atomicAdd(&cnt[g], seg_end - lane + 1);
    }
}

// sums -> means in place; inv = 1/count (0 when empty), as the reference multiplies.
__global__ void means_kernel(double* sy, double* sW, double* sZ, const int* __restrict__ cnt, int G, int K, int L) {
    const size_t total = (size_t)G * (1 + K + L);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        double* p;
        int g;
        if (e < (size_t)G) {
            g = (int)e;
            p = sy + e;
        } else if (e < (size_t)G * (1 + K)) {
            size_t o = e - G;
            g = (int)(o / K);
            p = sW + o;
        } else {
            size_t o = e - (size_t)G * (1 + K);
            g = (int)(o / L);
            p = sZ + o;
        }
        int c = cnt[g];
        double inv = c > 0 ? 1.0 / (double)c : 0.0;
        *p *= inv;
    }
}

__global__ void subtract_kernel(double* y, double* W, double* Z, const int* __restrict__ ids, size_t n, int K, int L,
                                size_t ld, const double* __restrict__ my, const double* __restrict__ mW,
                                const double* __restrict__ mZ) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int g = ids[i] - 1;
        if (g < 0) continue;
        y[i] -= my[g];
        const double* rw = mW + (size_t)g * K;
        for (int k = 0; k < K; ++k) W[i + (size_t)k * ld] -= rw[k];
        const double* rz = mZ + (size_t)g * L;
        for (int k = 0; k < L; ++k) Z[i + (size_t)k * ld] -= rz[k];
    }
}

__global__ void mix_kernel(double* mean, double* prev, size_t n, double r) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double p = prev[i];
        double v = p + r * (mean[i] - p);
        prev[i] = v;
        mean[i] = v;
    }
}

__global__ void axpy_kernel(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] += a * x[i];
}

// max |x| -> device slot (bits of a non-negative double order like unsigned integers)
__global__ __launch_bounds__(kBlock) void absmax_kernel(const double* __restrict__ x, size_t n,
                                                        unsigned long long* out) {
    __shared__ double part[kBlock / kWave];
    double m = 0.0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) m = fmax(m, fabs(x[i]));
    m = wave_max(m);
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / kWave; ++w) m = fmax(m, part[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

struct ColIndex {
    int src[256];
};

__global__ void copy_columns_kernel(const double* __restrict__ src, size_t ld_src, ColIndex idx, int n_idx,
                                    size_t n_rows, double* __restrict__ dst, size_t ld_dst, int dst_col0) {
    const int j = blockIdx.y;
    if (j >= n_idx) return;
    const double* s = src + (size_t)idx.src[j] * ld_src;
    double* d = dst + (size_t)(dst_col0 + j) * ld_dst;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_rows; i += stride) d[i] = s[i];
}

}  // namespace
}  // namespace fehip

using namespace fehip;

extern "C" {

int fe_gpu_fe_accumulate(const double* y, const double* W, const double* Z, const int* fe_ids, size_t n_obs,
                         int n_groups, int n_reg, int n_inst, size_t leading_dim, double* group_sum_y,
                         double* group_sum_W, double* group_sum_Z, int* group_counts) {
    if (n_obs == 0) return set_ok();
    if (n_groups <= 0) return set_ok();
    const int K = n_reg > 0 && W ? n_reg : 0;
    const int L = n_inst > 0 && Z && group_sum_Z ? n_inst : 0;
    const size_t lds = sizeof(double) * (size_t)(1 + K + L) * n_groups + sizeof(int) * (size_t)n_groups;
    if (lds <= 48 * 1024) {
        unsigned grid = stream_grid(n_obs, kBlock, 4);
        hipLaunchKernelGGL(accumulate_lds_kernel, dim3(grid), dim3(kBlock), lds, 0, y, W, Z, fe_ids, n_obs,
                           n_groups, K, L, leading_dim, group_sum_y, group_sum_W, group_sum_Z, group_counts);
    } else {
        unsigned grid = stream_grid(n_obs, kBlock, 8);
        hipLaunchKernelGGL(accumulate_runs_kernel, dim3(grid), dim3(kBlock), 0, 0, y, W, Z, fe_ids, n_obs, n_groups,
                           K, L, leading_dim, group_sum_y, group_sum_W, group_sum_Z, group_counts);
    }
    FEHIP_LAUNCH_CHECK("accumulate kernel");
    return set_ok();
}

int fe_gpu_fe_compute_means(double* group_sum_y, double* group_sum_W, double* group_sum_Z, const int* group_counts,
                            int n_groups, int n_reg, int n_inst) {
    if (n_groups <= 0) return set_ok();
    const int K = n_reg > 0 && group_sum_W ? n_reg : 0;
    const int L = n_inst > 0 && group_sum_Z ? n_inst : 0;
    size_t total = (size_t)n_groups * (1 + K + L);
    hipLaunchKernelGGL(means_kernel, dim3(stream_grid(total, kBlock)), dim3(kBlock), 0, 0, group_sum_y, group_sum_W,
                       group_sum_Z, group_counts, n_groups, K, L);
    FEHIP_LAUNCH_CHECK("means kernel");
    return set_ok();
}

int fe_gpu_fe_subtract(double* y, double* W, double* Z, const int* fe_ids, size_t n_obs, int n_reg, int n_inst,
                       size_t leading_dim, const double* group_mean_y, const double* group_mean_W,
                       const double* group_mean_Z, double relaxation) {
    (void)relaxation;  // unused by the reference kernel too
    if (n_obs == 0) return set_ok();
    const int K = n_reg > 0 && W ? n_reg : 0;
    const int L = n_inst > 0 && Z && group_mean_Z ? n_inst : 0;
    hipLaunchKernelGGL(subtract_kernel, dim3(stream_grid(n_obs, kBlock)), dim3(kBlock), 0, 0, y, W, Z, fe_ids, n_obs,
                       K, L, leading_dim, group_mean_y, group_mean_W, group_mean_Z);
    FEHIP_LAUNCH_CHECK("subtract kernel");
    return set_ok();
}

int fe_gpu_mix_means(double* mean, double* prev, size_t n, double relaxation) {
    if (n == 0) return set_ok();
    hipLaunchKernelGGL(mix_kernel, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, 0, mean, prev, n, relaxation);
    FEHIP_LAUNCH_CHECK("mix kernel");
    return set_ok();
}

int fe_gpu_axpy(int n, double alpha, const double* x, double* y) {
    if (n <= 0 || alpha == 0.0 || !x || !y) return set_ok();
    hipLaunchKernelGGL(axpy_kernel, dim3(stream_grid((size_t)n, kBlock)), dim3(kBlock), 0, 0, n, alpha, x, y);
    FEHIP_LAUNCH_CHECK("axpy kernel");
    return set_ok();
}

int fe_gpu_absmax(const double* data, long long n, double* out) {
    if (!data || !out || n <= 0) {
        if (out) *out = 0.0;
        return set_error("Invalid arguments to absmax");
    }
    auto* slot = static_cast<unsigned long long*>(workspace(sizeof(unsigned long long), 1));
    if (!slot) return set_error("absmax: workspace allocation failed");
    FEHIP_CHECK(hipMemsetAsync(slot, 0, sizeof(unsigned long long), 0), "absmax memset");
    hipLaunchKernelGGL(absmax_kernel, dim3(stream_grid((size_t)n, kBlock)), dim3(kBlock), 0, 0, data, (size_t)n,
                       slot);
    FEHIP_LAUNCH_CHECK("absmax kernel");
    unsigned long long bits = 0;
    FEHIP_CHECK(hipMemcpy(&bits, slot, sizeof(bits), hipMemcpyDeviceToHost), "absmax copy");
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *out = v;
    return set_ok();
}

int fe_gpu_copy_columns(const double* src, int ld_src, const int* indices, int n_indices, int n_rows, double* dst,
                        int ld_dst, int dest_offset) {
    if (n_indices <= 0 || n_rows <= 0) return set_ok();
    if (!src || !dst || !indices) return set_error("Invalid buffers for column copy");
    for (int j = 0; j < n_indices; ++j)
        if (indices[j] - 1 < 0) return set_error("Column index out of range");
    for (int j0 = 0; j0 < n_indices; j0 += 256) {
        ColIndex ci;
        int cnt = n_indices - j0 < 256 ? n_indices - j0 : 256;
        for (int j = 0; j < cnt; ++j) ci.src[j] = indices[j0 + j] - 1;
        unsigned gx = stream_grid((size_t)n_rows * cnt, kBlock, 8);
        gx = (gx + cnt - 1) / cnt;
        if (gx < 1) gx = 1;
        hipLaunchKernelGGL(copy_columns_kernel, dim3(gx, cnt), dim3(kBlock), 0, 0, src, (size_t)ld_src, ci, cnt,
                           (size_t)n_rows, dst, (size_t)ld_dst, dest_offset + j0);
        FEHIP_LAUNCH_CHECK("copy columns kernel");
    }
    return set_ok();
}

}  // extern "C"
