// Wavefront-level segmented scans for CDNA (64-wide waves): DPP row shifts and row broadcasts with
// lane predicates derived from head ballots.  Used by the operator kernels of the within transform
// and by the cluster-score kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fehip {

// x moved across lanes by a DPP control.  Every row is written (row_mask 0xf) and lanes without a
// source lane read 0 (bound_ctrl), so no old value is needed; all 64 lanes must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Lane predicates of a segmented inclusive scan whose segment of this lane starts at lane s.
struct ScanMask {
    bool m1, m2, m4, m8, b15, b31;
    __device__ __forceinline__ ScanMask(int lane, int s) {
        const int rl = lane & 15;
        m1 = rl >= 1 && lane - 1 >= s;
        m2 = rl >= 2 && lane - 2 >= s;
        m4 = rl >= 4 && lane - 4 >= s;
        m8 = rl >= 8 && lane - 8 >= s;
        b15 = (lane & 31) >= 16 && s < (lane & ~15);
        b31 = lane >= 32 && s < 32;
    }
};

// The same predicates as 0.0 / 1.0 factors: a step is then one fused multiply-add per column
// (values are finite; lanes without a DPP source read 0).
struct ScanFactors {
    double f1, f2, f4, f8, f15, f31;
    __device__ __forceinline__ explicit ScanFactors(const ScanMask& m)
        : f1(m.m1 ? 1.0 : 0.0), f2(m.m2 ? 1.0 : 0.0), f4(m.m4 ? 1.0 : 0.0), f8(m.m8 ? 1.0 : 0.0),
          f15(m.b15 ? 1.0 : 0.0), f31(m.b31 ? 1.0 : 0.0) {}
};

__device__ __forceinline__ double seg_scan_f(double x, const ScanFactors& f) {
    x = fma(dpp_move<0x111>(x), f.f1, x);  // row_shr:1
    x = fma(dpp_move<0x112>(x), f.f2, x);  // row_shr:2
    x = fma(dpp_move<0x114>(x), f.f4, x);  // row_shr:4
    x = fma(dpp_move<0x118>(x), f.f8, x);  // row_shr:8
    x = fma(dpp_move<0x142>(x), f.f15, x); // row_bcast:15
    x = fma(dpp_move<0x143>(x), f.f31, x); // row_bcast:31
    return x;
}

// seg_scan_f on N columns at once, step by step, so the lane moves of one column overlap the
// arithmetic of the others (no idle cycles between a write and the DPP read of it).
template <int N>
__device__ __forceinline__ void seg_scan_cols(double (&x)[N], const ScanFactors& f) {
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] = fma(dpp_move<0x111>(x[c]), f.f1, x[c]);
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] = fma(dpp_move<0x112>(x[c]), f.f2, x[c]);
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] = fma(dpp_move<0x114>(x[c]), f.f4, x[c]);
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] = fma(dpp_move<0x118>(x[c]), f.f8, x[c]);
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] = fma(dpp_move<0x142>(x[c]), f.f15, x[c]);
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] = fma(dpp_move<0x143>(x[c]), f.f31, x[c]);
}

// sum of x over lanes [s, lane] of the wave (every lane of the wave must take part)
__device__ __forceinline__ double seg_scan(double x, const ScanMask& m) {
    double t;
    t = dpp_move<0x111>(x);  // row_shr:1
    x += m.m1 ? t : 0.0;
    t = dpp_move<0x112>(x);  // row_shr:2
    x += m.m2 ? t : 0.0;
    t = dpp_move<0x114>(x);  // row_shr:4
    x += m.m4 ? t : 0.0;
    t = dpp_move<0x118>(x);  // row_shr:8
    x += m.m8 ? t : 0.0;
    t = dpp_move<0x142>(x);  // row_bcast:15
    x += m.b15 ? t : 0.0;
    t = dpp_move<0x143>(x);  // row_bcast:31
    x += m.b31 ? t : 0.0;
    return x;
}

__device__ __forceinline__ unsigned long long mask_le(int lane) { return (2ull << lane) - 1ull; }

}  // namespace fehip
