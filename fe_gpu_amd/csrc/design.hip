// fehip_design_*: categorical dummies, interactions and products of a formula, generated on the device
// from the raw columns.
//
// The reference builds the expanded design on the host (build_formula_design_matrix,
// src/src_main.f90:595-736): per categorical variable the levels are the first MAX_CATEGORIES = 40
// distinct values that lie within CATEGORY_TOL = 1e-6 of an integer, in row order, then sorted
// (collect_levels_for_column, :1389-1432); a categorical term gives one 0/1 column per level but the
// base, in ascending level order, set where nint(x) equals the level (append_main_term_columns,
// :1178-1230); an interaction is the column-wise product of its factors' blocks, the first factor's
// columns outermost (append_interaction_columns, :1276-1335).  For the wide 2SLS design that is some
// 40 GB of host memory and PCIe traffic before the solve; here only the raw columns are resident and
// every output value is written once, by a kernel, from a per-column list of factors.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/fehip.h"
#include "common.hpp"

namespace fehip {
namespace {

constexpr int kMaxLevels = 40;        // MAX_CATEGORIES (src_main.f90:15)
constexpr double kCategoryTol = 1e-6; // CATEGORY_TOL (src_main.f90:16)
constexpr int kLevelBlocks = 1024;

// Fortran nint: nearest integer, halves away from zero
__device__ __forceinline__ long long nint_ll(double x) { return (long long)(x >= 0.0 ? floor(x + 0.5) : ceil(x - 0.5)); }

// Per block: the distinct values within CATEGORY_TOL of an integer over its contiguous row range, in
// order of first appearance, at most kMaxLevels of them (one wave walks the range 64 rows at a time;
// the lanes of a step are rows in order).
__global__ __launch_bounds__(64) void block_levels_kernel(const double* __restrict__ col, int64_t n, int64_t per_block,
                                                           int* __restrict__ out, int* __restrict__ out_n) {
    __shared__ int found[kMaxLevels];
    __shared__ int nfound;
    const int lane = threadIdx.x;
    if (lane == 0) nfound = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * per_block, r1 = std::min<int64_t>(n, r0 + per_block);
    for (int64_t base = r0; base < r1 && nfound < kMaxLevels; base += 64) {
        const int64_t r = base + lane;
        bool ok = false;
        int v = 0;
        if (r < r1) {
            const double x = col[r];
            const long long q = nint_ll(x);
            ok = fabs(x - (double)q) <= kCategoryTol && q >= INT32_MIN && q <= INT32_MAX;
            v = (int)q;
        }
        // lanes whose value is not yet in the list, added in lane (row) order
        bool fresh = ok;
        for (int j = 0; j < nfound && fresh; ++j) fresh = found[j] != v;
        unsigned long long F = __ballot(fresh);
        while (F) {
            const int j = __ffsll((long long)F) - 1;
            F &= F - 1ull;
            const int vj = __shfl(v, j, 64);
            int nf = nfound;
            bool dup = false;
            for (int k = 0; k < nf; ++k) dup |= found[k] == vj;
            __syncthreads();
            if (!dup && nf < kMaxLevels && lane == 0) found[nf] = vj, nfound = nf + 1;
            __syncthreads();
        }
    }
    __syncthreads();
    if (lane < nfound) out[(int64_t)blockIdx.x * kMaxLevels + lane] = found[lane];
    if (lane == 0) out_n[blockIdx.x] = nfound;
}

struct ColSpec {
    int n;
    int col[FEHIP_DESIGN_MAX_FACTORS];
    int level[FEHIP_DESIGN_MAX_FACTORS];
    int cat[FEHIP_DESIGN_MAX_FACTORS];
};

// out[row + j * ld_out] = product over the column's factors of (categorical: nint(x) == level; else x).
// One thread per row writes the row of every column of the launch: the few raw values a row needs are
// read from memory once and then come from the cache, and each column's store of a wave is 64
// consecutive doubles.
__global__ __launch_bounds__(256) void expand_kernel(const double* __restrict__ raw, int64_t ld_raw, int64_t n,
                                                     const ColSpec* __restrict__ spec, int n_cols,
                                                     double* __restrict__ out, int64_t ld_out) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        for (int j = 0; j < n_cols; ++j) {
            const ColSpec& s = spec[j];  // uniform: scalar loads
            double v = 1.0;
            for (int f = 0; f < s.n; ++f) {
                const double x = raw[r + (int64_t)s.col[f] * ld_raw];
                v *= s.cat[f] ? (nint_ll(x) == (long long)s.level[f] ? 1.0 : 0.0) : x;
            }
            __builtin_nontemporal_store(v, &out[r + (int64_t)j * ld_out]);
        }
    }
}

}  // namespace
}  // namespace fehip

using namespace fehip;

extern "C" int fehip_design_levels(const double* d_col, int64_t n, int32_t* levels, int32_t* n_levels,
                                   int32_t* truncated) {
    (void)hipGetLastError();
    if (!d_col || !levels || !n_levels || n < 0) return set_error("fehip_design_levels: bad arguments");
    *n_levels = 0;
    if (truncated) *truncated = 0;
    if (n == 0) return set_ok();
    const int nb = (int)std::min<int64_t>(kLevelBlocks, (n + 63) / 64);
    const int64_t per = (n + nb - 1) / nb;
    int *d_out = nullptr, *d_n = nullptr;
    if (!(d_out = static_cast<int*>(cache_alloc(sizeof(int) * (size_t)nb * kMaxLevels)))) return set_error("levels buffer");
    if (!(d_n = static_cast<int*>(cache_alloc(sizeof(int) * (size_t)nb)))) {
        cache_free(d_out);
        return set_error("levels buffer");
    }
    hipLaunchKernelGGL(block_levels_kernel, dim3(nb), dim3(64), 0, 0, d_col, n, per, d_out, d_n);
    std::vector<int> h((size_t)nb * kMaxLevels), hn(nb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = copy_sync(h.data(), d_out, sizeof(int) * h.size(), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = copy_sync(hn.data(), d_n, sizeof(int) * hn.size(), hipMemcpyDeviceToHost);
    cache_free(d_out);
    cache_free(d_n);
    FEHIP_CHECK(e, "design levels");
    // the blocks' lists in block (row) order give the first kMaxLevels distinct values of the column
    std::vector<int> first;
    bool more = false;
    for (int b = 0; b < nb && !more; ++b)
        for (int k = 0; k < hn[b]; ++k) {
            const int v = h[(size_t)b * kMaxLevels + k];
            if (std::find(first.begin(), first.end(), v) != first.end()) continue;
            if ((int)first.size() == kMaxLevels) {
                more = true;
                break;
            }
            first.push_back(v);
        }
    std::sort(first.begin(), first.end());
    for (size_t k = 0; k < first.size(); ++k) levels[k] = first[k];
    *n_levels = (int32_t)first.size();
    // the reference warns when the list filled up (src_main.f90:1415-1430)
    if (truncated) *truncated = (int)first.size() == kMaxLevels;
    return set_ok();
}

extern "C" int fehip_design_expand(const double* d_raw, int64_t ld_raw, int n_raw, int64_t n,
                                   const fehip_design_column* cols, int n_cols, double* d_out, int64_t ld_out) {
    (void)hipGetLastError();
    if (!d_raw || !cols || !d_out || n < 0 || n_cols < 0 || n_raw < 0 || ld_raw < n || ld_out < n)
        return set_error("fehip_design_expand: bad arguments");
    if (n == 0 || n_cols == 0) return set_ok();
    std::vector<ColSpec> hs((size_t)n_cols);
    for (int j = 0; j < n_cols; ++j) {
        const fehip_design_column& c = cols[j];
        if (c.n_factors < 1 || c.n_factors > FEHIP_DESIGN_MAX_FACTORS)
            return set_error("fehip_design_expand: column %d has %d factors", j, c.n_factors);
        hs[j].n = c.n_factors;
        for (int f = 0; f < FEHIP_DESIGN_MAX_FACTORS; ++f) {
            hs[j].col[f] = f < c.n_factors ? c.col[f] : 0;
            hs[j].level[f] = f < c.n_factors ? c.level[f] : 0;
            hs[j].cat[f] = f < c.n_factors ? (c.categorical[f] != 0) : 0;
            if (f < c.n_factors && (c.col[f] < 0 || c.col[f] >= n_raw))
                return set_error("fehip_design_expand: column %d reads raw column %d of %d", j, c.col[f], n_raw);
        }
    }
    ColSpec* ds = static_cast<ColSpec*>(cache_alloc(sizeof(ColSpec) * (size_t)n_cols));
    if (!ds) return set_error("design spec buffer");
    hipError_t e = hipMemcpyAsync(ds, hs.data(), sizeof(ColSpec) * (size_t)n_cols, hipMemcpyHostToDevice, 0);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(expand_kernel, dim3(stream_grid((size_t)n, 256)), dim3(256), 0, 0, d_raw, ld_raw, n,
                           (const ColSpec*)ds, n_cols, d_out, ld_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(0);
    cache_free(ds);
    FEHIP_CHECK(e, "design expand");
    return set_ok();
}
