// Device builder of cluster-intersection ids (the drop-in for
// fe_gpu_build_multi_cluster_ids, src/gpu/fe_gpu_runtime.cu:700-886).  The numbering is the
// CPU builder's (src/core/cluster_id_builder.c:146-194): id = 1 + dense rank of the mixed-radix
// key, so the result is bit-identical to the CPU path the reference uses in standard mode.
//
// MI355X design: the key is packed into the fewest bits that hold the largest key (found in
// the key pass), the (key, row) pairs are radix-sorted over only those bits with rocPRIM's
// onesweep sort, and a head-flag scan plus one scatter yields the ids.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "../../include/fe_gpu_runtime.h"
#include "common.hpp"

namespace fehip {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxDims = 8;

struct DimPtrs {
    const int* p[kMaxDims];
    unsigned long long stride[kMaxDims];
};

__global__ __launch_bounds__(kBlock) void keys_kernel(DimPtrs dp, int nd, long long n, unsigned long long* keys,
                                                      int* rows, unsigned long long* kmax) {
    unsigned long long m = 0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        unsigned long long k = 0;
        for (int d = 0; d < nd; ++d) {
            int v = dp.p[d][i];
            if (v > 0) k += (unsigned long long)(v - 1) * dp.stride[d];
        }
        keys[i] = k;
        rows[i] = (int)i;
        m = k > m ? k : m;
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(m, off, kWave);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(kmax, m);
}

__global__ void head_flags_kernel(const unsigned long long* __restrict__ k, long long n, int* flags) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        flags[i] = (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
}

__global__ void scatter_ids_kernel(const int* __restrict__ rows, const int* __restrict__ ids_sorted, long long n,
                                   int* out) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[rows[i]] = ids_sorted[i];
}

}  // namespace

// Shared with the panel code: ids (1-based dense key rank) for arbitrary dims; optionally the
// sorted row order.  Returns 0 or a set_error status.
int build_intersection_ids(const int* const* dims, const unsigned long long* strides, int nd, long long n, int* out,
                           int* n_clusters, hipStream_t stream) {
    if (nd <= 0 || nd > kMaxDims) return set_error("cluster builder: %d dims (1..%d supported)", nd, kMaxDims);
    if (n > 0x7fffffffLL) return set_error("cluster builder: n_obs exceeds int32 row indices");
    DimPtrs dp{};
    for (int d = 0; d < nd; ++d) {
        dp.p[d] = dims[d];
        dp.stride[d] = strides[d];
    }
    const size_t N = (size_t)n;
    // workspace layout: keys[2N] u64 | rows[2N] i32 | flags[N] i32 | kmax u64 | rocprim temp
    size_t off_keys = 0;
    size_t off_rows = off_keys + 2 * N * sizeof(unsigned long long);
    size_t off_flags = off_rows + 2 * N * sizeof(int);
    size_t off_kmax = (off_flags + N * sizeof(int) + 255) / 256 * 256;
    size_t off_tmp = off_kmax + 256;
    size_t sort_tmp = 0, scan_tmp = 0;
    rocprim::double_buffer<unsigned long long> kb(nullptr, nullptr);
    rocprim::double_buffer<int> vb(nullptr, nullptr);
    FEHIP_CHECK(rocprim::radix_sort_pairs(nullptr, sort_tmp, kb, vb, N, 0, 64, stream), "radix sort size");
    FEHIP_CHECK(rocprim::inclusive_scan(nullptr, scan_tmp, (int*)nullptr, (int*)nullptr, N, rocprim::plus<int>(),
                                        stream),
                "scan size");
    size_t tmp = sort_tmp > scan_tmp ? sort_tmp : scan_tmp;
    char* ws = static_cast<char*>(workspace(off_tmp + tmp, 4));
    if (!ws) return set_error("cluster builder: workspace allocation failed");
    auto* k0 = reinterpret_cast<unsigned long long*>(ws + off_keys);
    auto* k1 = k0 + N;
    int* r0 = reinterpret_cast<int*>(ws + off_rows);
    int* r1 = r0 + N;
    int* flags = reinterpret_cast<int*>(ws + off_flags);
    auto* kmax = reinterpret_cast<unsigned long long*>(ws + off_kmax);
    void* temp = ws + off_tmp;

    FEHIP_CHECK(hipMemsetAsync(kmax, 0, sizeof(unsigned long long), stream), "kmax reset");
    hipLaunchKernelGGL(keys_kernel, dim3(stream_grid(N, kBlock)), dim3(kBlock), 0, stream, dp, nd, n, k0, r0, kmax);
    FEHIP_LAUNCH_CHECK("keys kernel");
    unsigned long long kmax_h = 0;
    FEHIP_CHECK(hipMemcpyAsync(&kmax_h, kmax, sizeof(kmax_h), hipMemcpyDeviceToHost, stream), "kmax copy");
    FEHIP_CHECK(hipStreamSynchronize(stream), "kmax sync");
    int bits = 0;
    while (bits < 64 && (kmax_h >> bits) != 0) ++bits;
    rocprim::double_buffer<unsigned long long> keys(k0, k1);
    rocprim::double_buffer<int> rows(r0, r1);
    if (bits > 0) {
        size_t t = sort_tmp;
        FEHIP_CHECK(rocprim::radix_sort_pairs(temp, t, keys, rows, N, 0, bits, stream), "radix sort");
    }
    hipLaunchKernelGGL(head_flags_kernel, dim3(stream_grid(N, kBlock)), dim3(kBlock), 0, stream, keys.current(), n,
                       flags);
    FEHIP_LAUNCH_CHECK("head flags kernel");
    // ids_sorted written into the spare key buffer (reinterpreted as int)
    int* ids_sorted = reinterpret_cast<int*>(keys.alternate());
    size_t t2 = scan_tmp;
    FEHIP_CHECK(rocprim::inclusive_scan(temp, t2, flags, ids_sorted, N, rocprim::plus<int>(), stream), "id scan");
    hipLaunchKernelGGL(scatter_ids_kernel, dim3(stream_grid(N, kBlock)), dim3(kBlock), 0, stream, rows.current(),
                       ids_sorted, n, out);
    FEHIP_LAUNCH_CHECK("scatter kernel");
    int last = 0;
    FEHIP_CHECK(hipMemcpyAsync(&last, ids_sorted + (N - 1), sizeof(int), hipMemcpyDeviceToHost, stream), "count copy");
    FEHIP_CHECK(hipStreamSynchronize(stream), "cluster builder sync");
    *n_clusters = last;
    return 0;
}

}  // namespace fehip

using namespace fehip;

extern "C" int fe_gpu_build_multi_cluster_ids(const void* const* fe_ptrs, const unsigned long long* strides,
                                              int n_dims, long long n_obs, int* out_ids, int* out_n_clusters) {
    if (!fe_ptrs || !strides || !out_ids || !out_n_clusters || n_dims <= 0)
        return set_error("Invalid arguments to GPU cluster builder");
    if (n_obs <= 0) {
        *out_n_clusters = 0;
        return set_ok();
    }
    int st = build_intersection_ids(reinterpret_cast<const int* const*>(fe_ptrs), strides, n_dims, n_obs, out_ids,
                                    out_n_clusters, 0);
    if (st != 0) return st;
    if (*out_n_clusters == 0) return set_error("GPU cluster builder produced zero clusters");
    return set_ok();
}
