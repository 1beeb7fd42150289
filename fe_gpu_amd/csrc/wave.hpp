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

namespace fehip {

// ---- FP32 segmented scans: one fused multiply-add with a DPP source per column and step ----
// x += dpp(x) * f in one v_fmac_f32_dpp (src0 moved across lanes, src1 the lane's 0/1 factor).  The
// backend does not fold a 32-bit DPP move into the multiply-add, so the step is written out here.  A
// DPP read of a VGPR needs two wait states after the VALU write of it: each block starts with
// s_nop 1, and inside a block every column is read once and written once, in order.  Lanes without a
// source lane keep their value (the factor is 0 there anyway); all 64 lanes must be active.
#define FEHIP_FMAC_DPP(i, f, ctrl) "v_fmac_f32_dpp %" #i ", %" #i ", %" #f " " ctrl " row_mask:0xf bank_mask:0xf\n\t"

#define FEHIP_SCAN_STEP_FNS(NAME, CTRL)                                                                              \
    __device__ __forceinline__ void NAME##4(float& a, float& b, float& c, float& d, float f) {                      \
        asm("s_nop 1\n\t" FEHIP_FMAC_DPP(0, 4, CTRL) FEHIP_FMAC_DPP(1, 4, CTRL) FEHIP_FMAC_DPP(2, 4, CTRL)          \
                FEHIP_FMAC_DPP(3, 4, CTRL)                                                                           \
            : "+v"(a), "+v"(b), "+v"(c), "+v"(d)                                                                     \
            : "v"(f));                                                                                               \
    }                                                                                                                \
    __device__ __forceinline__ void NAME##2(float& a, float& b, float f) {                                          \
        asm("s_nop 1\n\t" FEHIP_FMAC_DPP(0, 2, CTRL) FEHIP_FMAC_DPP(1, 2, CTRL) : "+v"(a), "+v"(b) : "v"(f));        \
    }                                                                                                                \
    __device__ __forceinline__ void NAME##1(float& a, float f) {                                                    \
        asm("s_nop 1\n\t" FEHIP_FMAC_DPP(0, 1, CTRL) : "+v"(a) : "v"(f));                                            \
    }

FEHIP_SCAN_STEP_FNS(scan_shr1_, "row_shr:1")
FEHIP_SCAN_STEP_FNS(scan_shr2_, "row_shr:2")
FEHIP_SCAN_STEP_FNS(scan_shr4_, "row_shr:4")
FEHIP_SCAN_STEP_FNS(scan_shr8_, "row_shr:8")
FEHIP_SCAN_STEP_FNS(scan_bc15_, "row_bcast:15")
FEHIP_SCAN_STEP_FNS(scan_bc31_, "row_bcast:31")
#undef FEHIP_SCAN_STEP_FNS
#undef FEHIP_FMAC_DPP

#define FEHIP_SCAN_COLS(NAME, F)                                         \
    {                                                                    \
        int c = 0;                                                       \
        for (; c + 4 <= N; c += 4) NAME##4(x[c], x[c + 1], x[c + 2], x[c + 3], F); \
        if (c + 2 <= N) {                                                \
            NAME##2(x[c], x[c + 1], F);                                  \
            c += 2;                                                      \
        }                                                                \
        if (c < N) NAME##1(x[c], F);                                     \
    }

struct ScanFactorsF {
    float f1, f2, f4, f8, f15, f31;
    __device__ __forceinline__ explicit ScanFactorsF(const ScanMask& m)
        : f1(m.m1 ? 1.0f : 0.0f), f2(m.m2 ? 1.0f : 0.0f), f4(m.m4 ? 1.0f : 0.0f), f8(m.m8 ? 1.0f : 0.0f),
          f15(m.b15 ? 1.0f : 0.0f), f31(m.b31 ? 1.0f : 0.0f) {}
};

// the segmented inclusive scan of seg_scan_cols on N FP32 columns
template <int N>
__device__ __forceinline__ void seg_scan_cols_f32(float (&x)[N], const ScanFactorsF& f) {
#pragma unroll
    for (int step = 0; step < 1; ++step) {
        FEHIP_SCAN_COLS(scan_shr1_, f.f1)
        FEHIP_SCAN_COLS(scan_shr2_, f.f2)
        FEHIP_SCAN_COLS(scan_shr4_, f.f4)
        FEHIP_SCAN_COLS(scan_shr8_, f.f8)
        FEHIP_SCAN_COLS(scan_bc15_, f.f15)
        FEHIP_SCAN_COLS(scan_bc31_, f.f31)
    }
}
#undef FEHIP_SCAN_COLS

}  // namespace fehip
