// fehip_fe_values: the estimated fixed effects of a fitted regression, per absorbed dim.
//
// Replaces reconstruct_and_save_fixed_effects (src/core/fe_pipeline.f90:1619-1715), which runs an
// alternating projection on the host over r = y - W beta: per dim, the group means of the current
// residual, centred by their row-weighted mean, are added to the dim's values and taken off the
// residual, until the largest update is below the tolerance.  Its fixed point is unique on a connected
// design: the fitted part D v = P_D r - mean(r) split into per-dim values that each have row-weighted
// mean zero.  Here the same fixed point comes from the FE-space solve the within-transform already
// runs: the tables of r over the non-primary dims (PCG with the primary eliminated), the primary's
// values as segment means of r - D_o alpha_o (the elimination's own formula), then each dim centred by
// its row-weighted mean.  Levels without rows get 0 and count 0 (the reference leaves them out of its
// CSV files, write_fe_outputs at :1717-1748).
#include <algorithm>
#include <vector>

#include "../../include/fehip.h"
#include "common.hpp"
#include "panel.hpp"
#include "within.hpp"

namespace fehip {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxDims = 8;

__global__ void raw_residual_kernel(const double* __restrict__ y, const double* __restrict__ W, int64_t ld, int K,
                                    const double* __restrict__ beta, int64_t n, double* __restrict__ r) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = y[i];
        for (int k = 0; k < K; ++k) v -= W[i + (int64_t)k * ld] * beta[k];
        r[i] = v;
    }
}

// the other absorbed dims of the primary's segment means: panel-order ids and table offsets
struct OtherDims {
    const int* ids[kMaxDims];
    int64_t off[kMaxDims];
    int n;
};

// primary value of segment g: mean over its panel rows of r - sum_o alpha_o[id_o] (one thread per segment)
__global__ void primary_values_kernel(const double* __restrict__ r, const int* __restrict__ perm,
                                      const int64_t* __restrict__ seg_off, int64_t G, OtherDims od,
                                      const double* __restrict__ alpha, double* __restrict__ val,
                                      int* __restrict__ cnt) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G; g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = seg_off[g], b = seg_off[g + 1];
        double s = 0.0;
        for (int64_t i = a; i < b; ++i) {
            double v = r[perm ? perm[i] : i];
            for (int o = 0; o < od.n; ++o) v -= alpha[od.off[o] + od.ids[o][i]];
            s += v;
        }
        val[g] = b > a ? s / (double)(b - a) : 0.0;
        cnt[g] = (int)(b - a);
    }
}

__global__ void copy_table_kernel(const double* __restrict__ src, int64_t G, double* __restrict__ dst) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G; g += (int64_t)gridDim.x * blockDim.x)
        dst[g] = src[g];
}

// per block: sum of cnt * val and of cnt (fixed order within the block)
__global__ __launch_bounds__(kBlock) void weighted_sum_kernel(const double* __restrict__ val, const int* __restrict__ cnt,
                                                              int64_t G, double* __restrict__ part) {
    __shared__ double s0[kBlock], s1[kBlock];
    double a = 0.0, c = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G; g += (int64_t)gridDim.x * blockDim.x) {
        a += (double)cnt[g] * val[g];
        c += (double)cnt[g];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = c;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

__global__ void center_kernel(double* __restrict__ val, const int* __restrict__ cnt, int64_t G, double c) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G; g += (int64_t)gridDim.x * blockDim.x)
        val[g] = cnt[g] > 0 ? val[g] - c : 0.0;
}

template <typename T>
struct Owned {
    T* p = nullptr;
    ~Owned() {
        if (p) cache_free(p);
    }
    int get(size_t count) {
        p = static_cast<T*>(cache_alloc(sizeof(T) * (count ? count : 1)));
        return p ? 0 : set_error("device allocation of %zu bytes failed", sizeof(T) * count);
    }
};

}  // namespace
}  // namespace fehip

using namespace fehip;

extern "C" int fehip_fe_values(fehip_panel* panel, const double* d_y, const double* d_W, int K, int64_t ld,
                               const double* beta, const fehip_demean_opts* opts, double* d_values, int32_t* d_counts,
                               fehip_demean_stats* stats) {
    (void)hipGetLastError();
    if (!panel || !d_y || !opts || !d_values || !d_counts || K < 0 || (K > 0 && (!d_W || !beta)))
        return set_error("fehip_fe_values: bad arguments");
    Panel* p = reinterpret_cast<Panel*>(panel);
    const int64_t n = p->n;
    if (ld <= 0) ld = n;
    const unsigned all = (1u << p->m) - 1u;
    const unsigned mask = (opts->fe_mask ? opts->fe_mask : all) & all;
    int e;
    // r = y - W beta (caller order)
    Owned<double> r, dbeta, u, alpha, part;
    if ((e = r.get((size_t)n)) || (e = dbeta.get((size_t)std::max(1, K))) || (e = u.get((size_t)n))) return e;
    if (K > 0) FEHIP_CHECK(copy_sync(dbeta.p, beta, sizeof(double) * K, hipMemcpyHostToDevice), "beta copy");
    hipLaunchKernelGGL(raw_residual_kernel, dim3(stream_grid((size_t)n, kBlock)), dim3(kBlock), 0, 0, d_y, d_W, ld, K,
                       (const double*)dbeta.p, n, r.p);
    FEHIP_LAUNCH_CHECK("raw residual");
    // tables of r over the absorbed non-primary dims
    const int64_t S = within_table_size(p, mask);
    if ((e = alpha.get((size_t)std::max<int64_t>(1, S)))) return e;
    WithinCols wc(1);
    wc.x[0] = r.p;
    wc.out[0] = u.p;
    wc.alpha[0] = alpha.p;
    fehip_demean_opts o = *opts;
    o.fe_mask = mask;
    o.precision = 0;
    WithinStats ws(1);
    if ((e = within_transform(p, wc, 1, nullptr, 1, o, 1, ws, 0))) return e;
    // per dim, in ascending dim order: values (tables, or the primary's segment means) and row counts
    OtherDims od{};
    {
        int64_t toff = 0;
        for (int d = 0; d < p->m; ++d) {
            if (!((mask >> d) & 1u) || d == p->P) continue;
            od.ids[od.n] = p->d[d].ids;
            od.off[od.n] = toff;
            ++od.n;
            toff += p->d[d].G;
        }
    }
    const unsigned rgrid = 256;
    if ((e = part.get((size_t)2 * rgrid))) return e;
    int64_t voff = 0, toff = 0;
    for (int d = 0; d < p->m; ++d) {
        if (!((mask >> d) & 1u)) continue;
        const int64_t G = p->d[d].G;
        double* val = d_values + voff;
        int* cnt = d_counts + voff;
        if (d == p->P) {
            hipLaunchKernelGGL(primary_values_kernel, dim3(stream_grid((size_t)G, kBlock)), dim3(kBlock), 0, 0,
                               (const double*)r.p, (const int*)p->perm, (const int64_t*)p->seg_off, G, od,
                               (const double*)alpha.p, val, cnt);
        } else {
            hipLaunchKernelGGL(copy_table_kernel, dim3(stream_grid((size_t)G, kBlock)), dim3(kBlock), 0, 0,
                               (const double*)alpha.p + toff, G, val);
            // rows per level: the panel's own counts
            FEHIP_CHECK(hipMemcpyAsync(cnt, p->d[d].counts, sizeof(int) * (size_t)G, hipMemcpyDeviceToDevice, 0),
                        "counts");
            toff += G;
        }
        FEHIP_LAUNCH_CHECK("fe values");
        // centre by the row-weighted mean (fe_pipeline.f90:1685-1704)
        hipLaunchKernelGGL(weighted_sum_kernel, dim3(rgrid), dim3(kBlock), 0, 0, (const double*)val, (const int*)cnt, G,
                           part.p);
        std::vector<double> h(2 * rgrid);
        FEHIP_CHECK(copy_sync(h.data(), part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost), "centre sums");
        double ws0 = 0.0, ws1 = 0.0;
        for (unsigned b = 0; b < rgrid; ++b) {
            ws0 += h[2 * b];
            ws1 += h[2 * b + 1];
        }
        const double c = ws1 > 0.0 ? ws0 / ws1 : 0.0;
        hipLaunchKernelGGL(center_kernel, dim3(stream_grid((size_t)G, kBlock)), dim3(kBlock), 0, 0, val, (const int*)cnt,
                           G, c);
        FEHIP_LAUNCH_CHECK("centre");
        voff += G;
    }
    FEHIP_CHECK(hipStreamSynchronize(0), "fe values sync");
    if (stats) {
        stats->iterations = ws.iterations;
        stats->converged = ws.converged;
        stats->max_update = ws.max_update;
        stats->seconds = ws.seconds;
        stats->bytes_moved = ws.bytes;
    }
    return set_ok();
}
