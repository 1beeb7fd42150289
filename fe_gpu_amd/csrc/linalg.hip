// Dense-algebra entry points of the drop-in C-ABI (SYRK, GEMV, GEMM, residual, dot,
// cluster scores and meat).  They replace the reference's cuBLAS wrappers and its
// cluster-score kernel (src/gpu/fe_gpu_linalg.cu:29-231) with own HIP kernels:
//   * every product with a long reduction over observations (W'W, W'y, Z'W, S'S) is a
//     split-K LDS-tiled kernel whose per-chunk partials are reduced in a fixed order, so
//     the results are deterministic run to run;
//   * the cluster scores merge runs of equal cluster ids inside each wavefront before the
//     FP64 atomic, so cluster-sorted rows pay one atomic per run and column.
#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/fe_gpu_runtime.h"
#include "common.hpp"
#include "kstats.hpp"

namespace fehip {
namespace {

constexpr int kBlock = 256;
constexpr int TM = 32;  // output tile rows
constexpr int TN = 32;  // output tile cols
constexpr int TK = 32;  // reduction slab

// C_tile (TM x TN) = sum_{p in [k0,k1)} op(A)[i][p] * op(B)[p][j]
// split==1: C = alpha*acc + beta*C directly; else partial[z] (m x n, col-major, ld m).
__global__ __launch_bounds__(kBlock) void gemm_tile_kernel(bool tA, bool tB, int m, int n, long long k,
                                                           const double* __restrict__ A, long long ldA,
                                                           const double* __restrict__ B, long long ldB,
                                                           long long kchunk, double alpha, double beta, double* C,
                                                           long long ldC, double* partial, bool upper_only) {
    __shared__ double As[TK][TM + 1];
    __shared__ double Bs[TK][TN + 1];
    const int i0 = blockIdx.x * TM;
    const int j0 = blockIdx.y * TN;
    if (upper_only && i0 > j0 + TN - 1) return;  // tile strictly below the diagonal
    const long long k0 = (long long)blockIdx.z * kchunk;
    long long k1 = k0 + kchunk;
    if (k1 > k) k1 = k;
    const int t = threadIdx.x;
    const int ti = t % 16, tj = t / 16;  // 2x2 outputs per thread: rows ti, ti+16; cols tj, tj+16
    double c00 = 0, c01 = 0, c10 = 0, c11 = 0;
    for (long long p0 = k0; p0 < k1; p0 += TK) {
        // load op(A)[i0..i0+TM)[p0..p0+TK) and op(B)[p0..p0+TK)[j0..j0+TN)
        for (int e = t; e < TM * TK; e += kBlock) {
            int pp, ii;
            if (tA) {  // A is k x m: consecutive threads walk p (contiguous in A)
                pp = e % TK;
                ii = e / TK;
            } else {   // A is m x k: consecutive threads walk i (contiguous)
                ii = e % TM;
                pp = e / TM;
            }
            long long p = p0 + pp;
            int i = i0 + ii;
            double v = 0.0;
            if (p < k1 && i < m) v = tA ? A[p + (long long)i * ldA] : A[i + p * ldA];
            As[pp][ii] = v;
        }
        for (int e = t; e < TN * TK; e += kBlock) {
            int pp, jj;
            if (!tB) {  // B is k x n: contiguous in p
                pp = e % TK;
                jj = e / TK;
            } else {    // B is n x k: contiguous in j
                jj = e % TN;
                pp = e / TN;
            }
            long long p = p0 + pp;
            int j = j0 + jj;
            double v = 0.0;
            if (p < k1 && j < n) v = tB ? B[j + p * ldB] : B[p + (long long)j * ldB];
            Bs[pp][jj] = v;
        }
        __syncthreads();
#pragma unroll 8
        for (int pp = 0; pp < TK; ++pp) {
            double a0 = As[pp][ti], a1 = As[pp][ti + 16];
            double b0 = Bs[pp][tj], b1 = Bs[pp][tj + 16];
            c00 = fma(a0, b0, c00);
            c01 = fma(a0, b1, c01);
            c10 = fma(a1, b0, c10);
            c11 = fma(a1, b1, c11);
        }
        __syncthreads();
    }
    const double acc[2][2] = {{c00, c01}, {c10, c11}};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            int i = i0 + ti + 16 * a, j = j0 + tj + 16 * b;
            if (i >= m || j >= n) continue;
            if (upper_only && i > j) continue;
            if (partial) {
                partial[(size_t)blockIdx.z * m * n + (size_t)j * m + i] = acc[a][b];
            } else {
                double* cp = C + i + (long long)j * ldC;
                *cp = beta == 0.0 ? alpha * acc[a][b] : alpha * acc[a][b] + beta * (*cp);
            }
        }
}

// C = A^T B for tall A (k x m) and B (k x n), column-major: the normal-equation products X'X, Z'Z, Z'X
// of wide designs, on the FP64 matrix cores (v_mfma_f64_16x16x4f64).  A workgroup of four waves owns a
// (32 NB) x (32 NB) output tile, each wave a (16 NB) x (16 NB) block of NB x NB MFMA tiles; slabs of
// kSlab rows of A and B are staged in LDS (loaded along the contiguous direction) and each lane feeds
// the matrix core its element (row = lane % 16 of the 16-block, k = lane / 16 of the 4-step).  Each
// split of the reduction writes its own partial tile; the splits are summed in a fixed order.  On the
// diagonal tiles of a symmetric product (A = B) one operand is loaded, and with only the upper
// triangle wanted the 16 x 16 blocks below the diagonal are skipped.
typedef double double4_t __attribute__((ext_vector_type(4)));

// Workgroups are dealt to the 8 XCDs round-robin by linear id; the grid is laid out so that all
// output tiles of one reduction split land on the same XCD one after another, and the row slabs they
// all read are shared through that XCD's L2.
constexpr int kXcds = 8;
constexpr int kSlab = 16;  // reduction rows per LDS slab

// RM: the reduction index runs along the leading dimension's complement, i.e. element (p, i) of the
// k x m operand is A[p ldA + i] (the A B^T shape of S'S for row-major cluster sums); otherwise A[p + i ldA].
// VEC: each load takes two adjacent elements along the contiguous direction (RM: m and n even;
// otherwise k even), so a pair is either wholly in range or wholly out.  Out-of-range positions read
// the operand's first element and are zeroed, without branches, so the loads of a slab stay in flight
// together.  Three (NB = 3) or four (NB = 2) workgroups per CU keep the matrix cores fed while others
// wait on their slabs.
template <int NB, bool RM, bool VEC>
__global__ __launch_bounds__(256, NB == 2 ? 4 : (NB == 3 ? 3 : 2)) void gemm_tn_mfma_kernel(
    int m, int n, long long k, const double* __restrict__ A, long long ldA, const double* __restrict__ B,
    long long ldB, int nsplit, int tiles_x, int tiles_y, double* partial, bool upper_only) {
    constexpr int BT = 32 * NB;            // block tile
    constexpr int KS = kSlab;
    constexpr int W = VEC ? 2 : 1;         // elements per load
    constexpr int PER = KS * BT / 256;     // elements of a slab each thread loads, per operand
    static_assert(PER % W == 0, "slab elements per thread");
    __shared__ double As[KS][BT + 1];
    __shared__ double Bs[KS][BT + 1];
    const int ntiles = tiles_x * tiles_y;
    const int xcd = blockIdx.x % kXcds, local = blockIdx.x / kXcds;
    const int split = xcd + kXcds * (local / ntiles), tile = local % ntiles;
    if (split >= nsplit) return;
    const int i0 = (tile % tiles_x) * BT, j0 = (tile / tiles_x) * BT;
    if (upper_only && i0 > j0 + BT - 1) return;
    // Split z takes the slabs z, z + nsplit, z + 2 nsplit, ... of the reduction: the workgroups in flight
    // read neighbouring rows, so a tall column-major operand's columns are read through a narrow window
    // of pages at a time (contiguous chunks per split would spread them over every page at once)
    const long long kstride = (long long)nsplit * KS;
    const long long k0 = (long long)split * KS;
    const int t = threadIdx.x, lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wi = (wave & 1) * 16 * NB, wj = (wave >> 1) * 16 * NB;
    const int lr = lane & 15, lk = lane >> 4;
    // the element(s) of operand X at reduction row p, column c (bounds lim, k1), zero outside
    auto fetch = [&](const double* X, long long ldX, int lim, long long p, long long k1, int c, double* dst) {
        const bool ok = p < k1 && c < lim;
        const long long off = ok ? (RM ? p * ldX + c : p + (long long)c * ldX) : 0;
        if constexpr (VEC) {
            const double2 v = *reinterpret_cast<const double2*>(X + off);
            dst[0] = ok ? v.x : 0.0;
            dst[1] = ok ? v.y : 0.0;
        } else {
            const double v = X[off];
            dst[0] = ok ? v : 0.0;
        }
    };
    auto body = [&](auto same_tag) {
        constexpr bool SAME = decltype(same_tag)::value;  // diagonal tile of A = B: Bs is As
        double4_t acc[NB][NB];
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) acc[a][b] = (double4_t){0.0, 0.0, 0.0, 0.0};
        // Software pipeline: the next slab is loaded into registers while the matrix cores work on the
        // current one from LDS, so the wait for global memory falls behind a slab of MFMAs.
        double ra[PER], rb[SAME ? 1 : PER];
        auto load_slab = [&](long long p0) {
            const long long k1 = p0 + KS < k ? p0 + KS : k;
#pragma unroll
            for (int q = 0; q < PER / W; ++q) {
                const int e = (t + 256 * q) * W;  // first of the W adjacent elements (same row / column)
                const int pp = RM ? e / BT : e % KS, ii = RM ? e % BT : e / KS;
                fetch(A, ldA, m, p0 + pp, k1, i0 + ii, ra + W * q);
                if constexpr (!SAME) fetch(B, ldB, n, p0 + pp, k1, j0 + ii, rb + W * q);
            }
        };
        auto store_slab = [&]() {
#pragma unroll
            for (int q = 0; q < PER / W; ++q) {
                const int e = (t + 256 * q) * W;
                const int pp = RM ? e / BT : e % KS, ii = RM ? e % BT : e / KS;
#pragma unroll
                for (int h = 0; h < W; ++h) {
                    const int ph = RM ? pp : pp + h, ih = RM ? ii + h : ii;
                    As[ph][ih] = ra[W * q + h];
                    if constexpr (!SAME) Bs[ph][ih] = rb[W * q + h];
                }
            }
        };
        if (k0 < k) load_slab(k0);
        for (long long p0 = k0; p0 < k; p0 += kstride) {
            store_slab();
            __syncthreads();
            if (p0 + kstride < k) load_slab(p0 + kstride);  // in flight during this slab's MFMAs
#pragma unroll
            for (int kk = 0; kk < KS / 4; ++kk) {
                double av[NB], bv[NB];
#pragma unroll
                for (int a = 0; a < NB; ++a) av[a] = As[kk * 4 + lk][wi + a * 16 + lr];
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    if constexpr (SAME)
                        bv[b] = As[kk * 4 + lk][wj + b * 16 + lr];
                    else
                        bv[b] = Bs[kk * 4 + lk][wj + b * 16 + lr];
                }
#pragma unroll
                for (int a = 0; a < NB; ++a)
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        if (SAME && upper_only && wi + a * 16 > wj + b * 16) continue;  // below the diagonal
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
                    }
            }
            __syncthreads();
        }
        // D[i][j] of a 16 x 16 tile: j = lane % 16, i = lane / 16 + 4 v (measured lane layout of the
        // f64 16x16x4 instruction on gfx950)
        double* out = partial + (size_t)split * m * n;
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = i0 + wi + a * 16 + lk + 4 * v, j = j0 + wj + b * 16 + lr;
                    if (i < m && j < n && !(upper_only && i > j)) out[(size_t)j * m + i] = acc[a][b][v];
                }
    };
    if (i0 == j0 && A == B && ldA == ldB)
        body(std::true_type{});
    else
        body(std::false_type{});
}

// C = alpha (sum of the split partials + ta tb') + beta C, where ta/tb (strides sa, sb; null when
// absent) are the reduction row a pair-loading kernel left out (an odd row count)
__global__ void gemm_finish_kernel(const double* __restrict__ partial, int nsplit, int m, int n, double alpha,
                                   double beta, double* C, long long ldC, bool upper_only, const double* ta,
                                   long long sa, const double* tb, long long sb) {
    const size_t total = (size_t)m * n;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        int i = (int)(e % m), j = (int)(e / m);
        if (upper_only && i > j) continue;
        double s = 0.0;
        for (int z = 0; z < nsplit; ++z) s += partial[(size_t)z * total + e];
        if (ta) s += ta[(long long)i * sa] * tb[(long long)j * sb];
        double* cp = C + i + (long long)j * ldC;
        *cp = beta == 0.0 ? alpha * s : alpha * s + beta * (*cp);
    }
}

}  // namespace

namespace {
// resident workgroups of one GEMM kernel instance on the device (its occupancy times the CUs)
unsigned resident_slots(int nb, bool rm, bool vec) {
    const size_t big = (size_t)1 << 40;
#define FEHIP_SLOTS(NBV, RMV, VECV) resident_grid(gemm_tn_mfma_kernel<NBV, RMV, VECV>, 256, 0, big)
    if (nb == 4) return rm ? (vec ? FEHIP_SLOTS(4, true, true) : FEHIP_SLOTS(4, true, false))
                           : (vec ? FEHIP_SLOTS(4, false, true) : FEHIP_SLOTS(4, false, false));
    if (nb == 3) return rm ? (vec ? FEHIP_SLOTS(3, true, true) : FEHIP_SLOTS(3, true, false))
                           : (vec ? FEHIP_SLOTS(3, false, true) : FEHIP_SLOTS(3, false, false));
    return rm ? (vec ? FEHIP_SLOTS(2, true, true) : FEHIP_SLOTS(2, true, false))
              : (vec ? FEHIP_SLOTS(2, false, true) : FEHIP_SLOTS(2, false, false));
#undef FEHIP_SLOTS
}

// A^T B on the matrix cores: tile size with the least padding of m and n (64, 96 or 128), split-K
// over the rows to fill the device.  The split-K partials come from the library's workspace, or from
// `fixed` (capacity `fixed_cap` doubles) when the caller must not allocate; the split count is then
// bounded by that capacity.
int gemm_tn_mfma(bool rm, int m, int n, long long k, double alpha, const double* A, long long ldA, const double* B,
                 long long ldB, double beta, double* C, long long ldC, bool upper_only, double* fixed,
                 size_t fixed_cap) {
    int nb = 2;
    double best = 0.0;
    for (int cand : {4, 3, 2}) {
        const int bt = 32 * cand;
        const double eff = (double)m * n / ((double)((m + bt - 1) / bt * bt) * ((n + bt - 1) / bt * bt));
        if (eff > best + 0.02) {
            best = eff;
            nb = cand;
        }
    }
    const int bt = 32 * nb;
    const int tx0 = (m + bt - 1) / bt, ty0 = (n + bt - 1) / bt;
    const long long tiles = (long long)tx0 * ty0;
    // tiles that do work (the upper triangle of a symmetric product) times the splits fill the device's
    // resident workgroup slots in whole rounds, so no round runs partly empty
    long long active = tiles;
    if (upper_only)
        for (int a = 0; a < tx0; ++a)
            for (int b = 0; b < ty0; ++b) active -= (a * bt > b * bt + bt - 1) ? 1 : 0;
    // pair loads: even leading dims, 16-byte aligned operands, and pairs that never straddle an edge
    // (RM: even m and n; otherwise an even row count, an odd last row being added by the finish pass)
    const bool vec = (ldA % 2 == 0) && (ldB % 2 == 0) && ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) &&
                     (!rm || (m % 2 == 0 && n % 2 == 0)) && (rm || k >= 2);
    const bool tail = vec && !rm && (k % 2 == 1);
    const long long kk = tail ? k - 1 : k;  // rows the kernel reduces
    // splits are dealt to the XCDs round-robin with all their tiles, so a multiple of the XCD count; the
    // active workgroups fill each XCD's slots about sixteen times over: diagonal tiles of a symmetric
    // product finish sooner than the others, and the many short rounds even out the slots' finishing
    // times (the partials stay within 1 GiB; the finish pass reads them once)
    const long long slots = (long long)resident_slots(nb, rm, vec);
    long long nsplit = kXcds * std::max<long long>(1, (16 * slots) / (kXcds * std::max<long long>(1, active)));
    nsplit = std::min<long long>(nsplit, std::max<long long>(1, kk / 1024));
    nsplit = std::min<long long>(nsplit, std::max<long long>(1, (1024LL << 20) / ((long long)m * n * 8)));
    if (fixed) nsplit = std::min<long long>(nsplit, std::max<long long>(1, (long long)(fixed_cap / ((size_t)m * n))));
    // every split must own at least one slab (the slabs are interleaved across the splits)
    nsplit = std::max<long long>(1, std::min<long long>(nsplit, (kk + kSlab - 1) / kSlab));
    double* partial = fixed ? fixed : static_cast<double*>(workspace((size_t)nsplit * m * n * sizeof(double), 2));
    if (!partial) return set_error("gemm: workspace allocation failed");
    const int tx = (m + bt - 1) / bt, ty = (n + bt - 1) / bt;
    const long long per_xcd = (nsplit + kXcds - 1) / kXcds;
    const unsigned grid = (unsigned)(kXcds * per_xcd * tx * ty);
#define FEHIP_MFMA_LAUNCH(NBV, RMV, VECV)                                                                        \
    hipLaunchKernelGGL((gemm_tn_mfma_kernel<NBV, RMV, VECV>), dim3(grid), dim3(256), 0, 0, m, n, kk, A, ldA, B, ldB, \
                       (int)nsplit, tx, ty, partial, upper_only)
#define FEHIP_MFMA_RM(NBV)                                              \
    do {                                                                \
        if (rm) {                                                       \
            if (vec) FEHIP_MFMA_LAUNCH(NBV, true, true);                 \
            else FEHIP_MFMA_LAUNCH(NBV, true, false);                    \
        } else {                                                        \
            if (vec) FEHIP_MFMA_LAUNCH(NBV, false, true);                \
            else FEHIP_MFMA_LAUNCH(NBV, false, false);                   \
        }                                                               \
    } while (0)
    if (nb == 4) FEHIP_MFMA_RM(4);
    else if (nb == 3) FEHIP_MFMA_RM(3);
    else FEHIP_MFMA_RM(2);
#undef FEHIP_MFMA_RM
#undef FEHIP_MFMA_LAUNCH
    FEHIP_LAUNCH_CHECK("gemm mfma kernel");
    hipLaunchKernelGGL(gemm_finish_kernel, dim3(stream_grid((size_t)m * n, kBlock)), dim3(kBlock), 0, 0, partial,
                       (int)nsplit, m, n, alpha, beta, C, ldC, upper_only, tail ? A + (k - 1) : nullptr, ldA,
                       tail ? B + (k - 1) : nullptr, ldB);
    FEHIP_LAUNCH_CHECK("gemm finish kernel");
    return set_ok();
}
}  // namespace

namespace {
int gemm_dispatch(bool tA, bool tB, int m, int n, long long k, double alpha, const double* A, long long ldA,
                  const double* B, long long ldB, double beta, double* C, long long ldC, bool upper_only,
                  double* fixed, size_t fixed_cap) {
    if (m <= 0 || n <= 0) return set_ok();
    // wide cross products over many observations go to the matrix cores
    if (tA != tB && m >= 16 && n >= 16 && k >= 4096) {  // A^T B (column-major) or A B^T (row-major view)
        const bool timed = !fixed && kstats_enabled();  // the timing events are not for the no-allocation path
        if (timed) kstat_begin(kStatGram, 0);
        const int e = gemm_tn_mfma(!tA, m, n, k, alpha, A, ldA, B, ldB, beta, C, ldC, upper_only, fixed, fixed_cap);
        if (timed) {
            const double pairs = upper_only ? 0.5 * (double)m * (m + 1) : (double)m * n;
            kstat_end(kStatGram, 0, 8.0 * (double)k * (upper_only ? m : m + n), 2.0 * pairs * (double)k);
        }
        return e;
    }
    const long long tiles = (long long)((m + TM - 1) / TM) * ((n + TN - 1) / TN);
    const long long target = 4LL * num_cus();
    long long nsplit = 1;
    if (k > 0 && tiles < target) {
        nsplit = target / tiles;
        long long max_by_k = (k + 255) / 256;
        if (nsplit > max_by_k) nsplit = max_by_k;
        long long max_by_mem = (256LL << 20) / ((long long)m * n * 8);
        if (nsplit > max_by_mem) nsplit = max_by_mem;
        if (fixed && nsplit > (long long)(fixed_cap / ((size_t)m * n))) nsplit = (long long)(fixed_cap / ((size_t)m * n));
        if (nsplit < 1) nsplit = 1;
        if (nsplit > 65535) nsplit = 65535;
    }
    long long kchunk = k > 0 ? (k + nsplit - 1) / nsplit : 1;
    kchunk = (kchunk + TK - 1) / TK * TK;
    if (k > 0) nsplit = (k + kchunk - 1) / kchunk;
    if (nsplit < 1) nsplit = 1;
    dim3 grid((m + TM - 1) / TM, (n + TN - 1) / TN, (unsigned)nsplit);
    if (nsplit == 1) {
        hipLaunchKernelGGL(gemm_tile_kernel, grid, dim3(kBlock), 0, 0, tA, tB, m, n, k, A, ldA, B, ldB, kchunk,
                           alpha, beta, C, ldC, (double*)nullptr, upper_only);
        FEHIP_LAUNCH_CHECK("gemm tile kernel");
        return set_ok();
    }
    double* partial = fixed ? fixed : static_cast<double*>(workspace((size_t)nsplit * m * n * sizeof(double), 2));
    if (!partial) return set_error("gemm: workspace allocation failed");
    hipLaunchKernelGGL(gemm_tile_kernel, grid, dim3(kBlock), 0, 0, tA, tB, m, n, k, A, ldA, B, ldB, kchunk, alpha,
                       beta, C, ldC, partial, upper_only);
    FEHIP_LAUNCH_CHECK("gemm tile kernel");
    hipLaunchKernelGGL(gemm_finish_kernel, dim3(stream_grid((size_t)m * n, kBlock)), dim3(kBlock), 0, 0, partial,
                       (int)nsplit, m, n, alpha, beta, C, ldC, upper_only, (const double*)nullptr, 0LL,
                       (const double*)nullptr, 0LL);
    FEHIP_LAUNCH_CHECK("gemm finish kernel");
    return set_ok();
}
}  // namespace

int gemm_impl(bool tA, bool tB, int m, int n, long long k, double alpha, const double* A, long long ldA,
              const double* B, long long ldB, double beta, double* C, long long ldC, bool upper_only) {
    return gemm_dispatch(tA, tB, m, n, k, alpha, A, ldA, B, ldB, beta, C, ldC, upper_only, nullptr, 0);
}

// ------------------------------------------------------------------------------------------------
// Scratch of the allocation-free SYRK: the split-K partials live in this fixed array, one copy per
// device for the life of the process.  Host code never passes the array's name to a kernel (on the
// host it is a stand-in, not a device address): it uses the address hipGetSymbolAddress returns for
// the current device, looked up once per device (at fe_gpu_linalg_init or the first call).
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

std::mutex g_scratch_mutex;
std::vector<double*> g_scratch_ptr;  // per device: the device address of `scratch`, or null

// the current device's address of `scratch`; nullptr (with the error message set) on failure
double* noalloc_scratch() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
        set_error("syrk_noalloc: no current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("syrk_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

namespace {

__global__ void residual_kernel(int n_rows, int n_cols, const double* __restrict__ W, long long ldW,
                                const double* __restrict__ beta, const double* __restrict__ y, double* r) {
    extern __shared__ double sb[];
    for (int c = threadIdx.x; c < n_cols; c += blockDim.x) sb[c] = beta[c];
    __syncthreads();
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_rows; i += stride) {
        double s = y[i];
        for (int c = 0; c < n_cols; ++c) s -= W[i + c * ldW] * sb[c];
        r[i] = s;
    }
}

__global__ __launch_bounds__(kBlock) void dot_partial_kernel(int n, const double* __restrict__ x,
                                                             const double* __restrict__ y, double* part) {
    __shared__ double ws[kBlock / kWave];
    double s = 0.0;
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) s = fma(x[i], y[i], s);
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += ws[w];
        part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBlock) void sum_partials_kernel(const double* __restrict__ part, int np, double* out) {
    __shared__ double ws[kBlock / kWave];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += ws[w];
        *out = t;
    }
}

__device__ __forceinline__ double run_sum(double v, int lane, int seg_end) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        double t = __shfl_down(v, off, kWave);
        if (lane + off <= seg_end) v += t;
    }
    return v;
}

__global__ __launch_bounds__(kBlock) void cluster_scores_kernel(const double* __restrict__ r,
                                                                const double* __restrict__ W,
                                                                const int* __restrict__ ids, long long n, int K,
                                                                long long ldW, int G, double* __restrict__ S) {
    const int lane = threadIdx.x & (kWave - 1);
    const long long wstride = (long long)gridDim.x * blockDim.x;
    for (long long base = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~(kWave - 1)); base < n;
         base += wstride) {
        const long long i = base + lane;
        int c = i < n ? ids[i] - 1 : -1;
        if (c < 0 || c >= G) c = -1 - lane;
        int cn = __shfl_down(c, 1, kWave), cp = __shfl_up(c, 1, kWave);
        bool tail = (lane == kWave - 1) || cn != c;
        bool head = (lane == 0) || cp != c;
        unsigned long long above = __ballot(tail) & (~0ull << lane);
        int seg_end = __ffsll((long long)above) - 1;
        const double ri = c >= 0 ? r[i] : 0.0;
        for (int j = 0; j < K; ++j) {
            double v = run_sum(c >= 0 ? ri * W[i + j * ldW] : 0.0, lane, seg_end);
            if (head && c >= 0) unsafeAtomicAdd(&S[c + (long long)j * G], v);
        }
    }
}

}  // namespace
}  // namespace fehip

using namespace fehip;

extern "C" {

int fe_gpu_linalg_init(void) {
    // look up the no-allocation scratch of the current device now, so its first call does no set-up
    if (!noalloc_scratch()) return 1;
    return set_ok();
}

int fe_gpu_linalg_shutdown(void) { return 0; }

int fe_gpu_syrk(int n_rows, int n_cols, double alpha, const double* W, int ldW, double beta, double* Q) {
    if (n_cols == 0) return set_ok();
    if (n_cols < 0 || n_rows < 0) return set_error("syrk: negative dimension");
    return gemm_impl(true, false, n_cols, n_cols, n_rows, alpha, W, ldW, W, ldW, beta, Q, n_cols, true);
}

int fe_gpu_syrk_noalloc(int n_rows, int n_cols, double alpha, const double* W, int ldW, double beta, double* Q) {
    if (n_cols == 0) return set_ok();
    if (n_cols < 0 || n_rows < 0) return set_error("syrk_noalloc: negative dimension");
    const size_t cap = kNoAllocScratchBytes / sizeof(double);
    if ((size_t)n_cols * (size_t)n_cols > cap)
        return set_error("syrk_noalloc: n_cols = %d needs %zu bytes of scratch, above the %zu-byte maximum (n_cols <= %d)",
                         n_cols, (size_t)n_cols * n_cols * sizeof(double), kNoAllocScratchBytes, 2048);
    double* fixed = noalloc_scratch();
    if (!fixed) return 1;
    return gemm_dispatch(true, false, n_cols, n_cols, n_rows, alpha, W, ldW, W, ldW, beta, Q, n_cols, true, fixed, cap);
}

int fe_gpu_gemv(int n_rows, int n_cols, double alpha, const double* W, int ldW, const double* y, double beta,
                double* b) {
    if (n_cols == 0) return set_ok();
    if (n_cols < 0 || n_rows < 0) return set_error("gemv: negative dimension");
    return gemm_impl(true, false, n_cols, 1, n_rows, alpha, W, ldW, y, n_rows > 0 ? n_rows : 1, beta, b, n_cols,
                     false);
}

int fe_gpu_residual(int n_rows, int n_cols, const double* W, int ldW, const double* beta, const double* y,
                    double* residual) {
    if (n_cols == 0) return set_ok();
    if (n_rows <= 0) return set_ok();
    hipLaunchKernelGGL(residual_kernel, dim3(stream_grid((size_t)n_rows, kBlock)), dim3(kBlock),
                       sizeof(double) * (size_t)n_cols, 0, n_rows, n_cols, W, (long long)ldW, beta, y, residual);
    FEHIP_LAUNCH_CHECK("residual kernel");
    return set_ok();
}

int fe_gpu_dot(int n_rows, const double* x, const double* y, double* result) {
    if (!result) return set_error("dot: null result");
    if (n_rows <= 0) {
        *result = 0.0;
        return set_ok();
    }
    unsigned grid = stream_grid((size_t)n_rows, kBlock, 4);
    double* part = static_cast<double*>(workspace(sizeof(double) * (grid + 1), 3));
    if (!part) return set_error("dot: workspace allocation failed");
    hipLaunchKernelGGL(dot_partial_kernel, dim3(grid), dim3(kBlock), 0, 0, n_rows, x, y, part);
    FEHIP_LAUNCH_CHECK("dot kernel");
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kBlock), 0, 0, part, (int)grid, part + grid);
    FEHIP_LAUNCH_CHECK("dot reduce kernel");
    FEHIP_CHECK(copy_sync(result, part + grid, sizeof(double), hipMemcpyDeviceToHost), "dot copy");
    return set_ok();
}

int fe_gpu_gemm(char transA, char transB, int m, int n, int k, double alpha, const double* A, int ldA,
                const double* B, int ldB, double beta, double* C, int ldC) {
    bool tA = !(transA == 'N' || transA == 'n');
    bool tB = !(transB == 'N' || transB == 'n');
    if (m < 0 || n < 0 || k < 0) return set_error("gemm: negative dimension");
    return gemm_impl(tA, tB, m, n, k, alpha, A, ldA, B, ldB, beta, C, ldC, false);
}

int fe_gpu_cluster_scores(const double* residual, const double* W, const int* cluster_ids, int n_rows, int n_cols,
                          int ldW, int n_clusters, double* scores) {
    if (n_cols == 0 || n_rows <= 0 || n_clusters <= 0) return set_ok();
    hipLaunchKernelGGL(cluster_scores_kernel, dim3(stream_grid((size_t)n_rows, kBlock, 8)), dim3(kBlock), 0, 0,
                       residual, W, cluster_ids, (long long)n_rows, n_cols, (long long)ldW, n_clusters, scores);
    FEHIP_LAUNCH_CHECK("cluster scores kernel");
    FEHIP_CHECK(hipDeviceSynchronize(), "cluster scores sync");
    return set_ok();
}

int fe_gpu_cluster_meat(int n_clusters, int n_cols, const double* scores, int ldScores, double* meat) {
    return fe_gpu_syrk(n_clusters, n_cols, 1.0, scores, ldScores, 0.0, meat);
}

}  // extern "C"
