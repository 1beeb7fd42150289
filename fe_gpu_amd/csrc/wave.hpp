// Wavefront-level segmented scans for CDNA (64-wide waves): DPP row shifts and row broadcasts with
// lane predicates derived from head ballots.  Used by the operator kernels of the within transform
// and by the cluster-score kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fehip {

// x moved across lanes by a DPP control (lanes without a source read 0)
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
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
