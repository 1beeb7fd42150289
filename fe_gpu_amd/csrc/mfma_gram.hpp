// Cross products (Gram matrices) of per-row values on the FP64 matrix cores.
//
// v_mfma_f64_16x16x4f64 computes D[i][j] += sum_k A[i][k] B[k][j] over four rows k, with A[i][k] taken
// from lane i + 16k and B[k][j] from lane j + 16k, and D[i][j] held by lane j + 16 (i % 4) as element
// i / 4 of its four-double accumulator.  For V'V, A[i][k] = B[k][i] = V[r0 + k][i], so both operands are
// the same value, V[r0 + lane / 16][lane % 16]: with a wave's 64 rows staged (row-major) in LDS, the
// update of up to 16 columns is sixteen instructions and the accumulator costs 8 registers, where
// per-lane FMAs would hold one register pair per (i <= j) pair.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace fehip {

typedef double gram4_t __attribute__((ext_vector_type(4)));

// doubles of LDS per wave: 64 staged rows of CT values, or one 16 x 16 record
template <int CT>
constexpr int gram_stage() {
    return 64 * CT > 256 ? 64 * CT : 256;
}

// one wave: V'V of the 64 rows st[row * CT + col] (CT <= 16) added to acc.  With CT <= 8 (CT <= 4) one
// instruction serves two (four) groups of four rows, each group in its own diagonal block of D (the
// off-diagonal blocks are discarded), so 8 (4) instructions cover the 64 rows instead of 16.
template <int CT>
constexpr int gram_block() {
    return CT <= 4 ? 4 : (CT <= 8 ? 8 : 16);
}

template <int CT>
__device__ __forceinline__ void gram_rows64(const double* st, int lane, gram4_t& acc) {
    static_assert(CT <= 16, "at most 16 columns per accumulator");
    constexpr int B = gram_block<CT>(), G = 16 / B;  // block width, row groups per instruction
    const int col = lane & 15, rr = lane >> 4;
    const int g = col / B, c = col - g * B;
#pragma unroll
    for (int k = 0; k < 16 / G; ++k) {
        const int row = 4 * G * k + 4 * g + rr;
        const double x = c < CT ? st[row * CT + c] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(x, x, acc, 0, 0, 0);
    }
}

// one wave: stage this lane's row (CT values, widened to FP64) and add the wave's V'V to acc
template <int CT, typename T>
__device__ __forceinline__ void gram_wave(double* st, const T (&v)[CT], int lane, gram4_t& acc) {
#pragma unroll
    for (int c = 0; c < CT; ++c) st[lane * CT + c] = (double)v[c];
    __builtin_amdgcn_wave_barrier();
    gram_rows64<CT>(st, lane, acc);
    __builtin_amdgcn_wave_barrier();
}

// gram_wave over the lanes where `on` holds only (the others' rows are zero): those rows are packed to the front
// of the staging area, the rows after them up to a whole instruction are cleared, and the products take
// ceil(rows / (4 * row groups)) instructions instead of 16 / row groups.  A subset whose clusters end at a few
// lanes of the tile (one per segment or run) then costs one or two instructions.
template <int CT, typename T>
__device__ __forceinline__ void gram_wave_sparse(double* st, const T (&v)[CT], bool on, int lane, gram4_t& acc) {
    constexpr int B = gram_block<CT>(), G = 16 / B, R = 4 * G;  // block width, row groups, rows per instruction
    const unsigned long long m = __ballot(on);
    const int n = __popcll(m);
    if (n == 0) return;  // uniform across the wave
    const int nr = (n + R - 1) / R * R;
    if (on) {
        const int pos = __popcll(m & ((1ull << lane) - 1ull));
#pragma unroll
        for (int c = 0; c < CT; ++c) st[pos * CT + c] = (double)v[c];
    }
    if (lane < nr - n) {
#pragma unroll
        for (int c = 0; c < CT; ++c) st[(n + lane) * CT + c] = 0.0;
    }
    __builtin_amdgcn_wave_barrier();
    const int col = lane & 15, rr = lane >> 4;
    const int g = col / B, c = col - g * B;
    for (int k = 0; k < nr / R; ++k) {
        const int row = R * k + 4 * g + rr;
        const double x = c < CT ? st[row * CT + c] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(x, x, acc, 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
}

// The block's waves add their accumulators in a fixed order and the diagonal blocks are folded; the
// CT x CT result is left in stage_all[i * 16 + j] (i, j < CT) and visible to every thread.  stage_all
// holds blockDim.x / 64 records of `stage` doubles.
template <int CT>
__device__ __forceinline__ void gram_block_sum(double* stage_all, int stage, const gram4_t& acc) {
    constexpr int B = gram_block<CT>(), G = 16 / B;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    __syncthreads();
    double* mine = stage_all + (size_t)wv * stage;
#pragma unroll
    for (int v = 0; v < 4; ++v) mine[(4 * v + (lane >> 4)) * 16 + (lane & 15)] = acc[v];
    __syncthreads();
    const int e = threadIdx.x;
    double t = 0.0;
    if (e < 256) {
        const int i = e >> 4, j = e & 15;
        if (i < B && j < B)
            for (int gg = 0; gg < G; ++gg) {
                const int f = (gg * B + i) * 16 + gg * B + j;
                for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += stage_all[(size_t)w * stage + f];
            }
    }
    __syncthreads();
    if (e < 256) stage_all[e] = t;
    __syncthreads();
}

// the block's record of 256 doubles (the CT x CT corner is meaningful)
template <int CT>
__device__ __forceinline__ void gram_store_block(double* stage_all, int stage, const gram4_t& acc, double* rec) {
    gram_block_sum<CT>(stage_all, stage, acc);
    for (int e = threadIdx.x; e < 256; e += blockDim.x) rec[e] = stage_all[e];
}

// the block's upper triangle (i <= j < KT, row by row) as KT (KT + 1) / 2 doubles
template <int KT>
__device__ __forceinline__ void gram_store_upper(double* stage_all, int stage, const gram4_t& acc, double* out) {
    gram_block_sum<KT>(stage_all, stage, acc);
    for (int e = threadIdx.x; e < KT * (KT + 1) / 2; e += blockDim.x) {
        int i = 0, rem = e;
        while (rem >= KT - i) {
            rem -= KT - i;
            ++i;
        }
        out[e] = stage_all[i * 16 + i + rem];
    }
}

}  // namespace fehip
