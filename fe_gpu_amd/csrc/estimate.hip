// fehip_estimate: the whole estimation core on a device-resident panel.
//
// Same conventions as the reference's GPU path, fe_gpu_estimate (src/core/fe_pipeline.f90:516-1896):
//   demean y, W, Z together (:616)  ->  one Gram pass [y W Z]'[y W Z] (replaces the separate
//   syrk / gemv / gemm calls of :647, :963-964) -> host selection of endogenous/exogenous/
//   instrument columns (:664-753), IV normal equations (:947-1032), zero-variance column filter
//   and Cholesky (:1034-1107) -> residual on W~ (:1168) -> homoskedastic or CGM sandwich over the
//   2^m-1 cluster subsets with the nested small-sample scalar and PSD clip (:1109-1420) ->
//   regression statistics (finalize_regression_stats).
// Cluster scores use the panel's cached row partitions: S_c = sum_{i in c} r_i x_i is a
// reduce-by-key over rows grouped by cluster (no atomics), and the meat sum_c S_c S_c' is
// accumulated in registers (narrow designs) or by a split-K Gram (wide designs).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "dense.hpp"
#include "panel.hpp"
#include "estimate.hpp"
#include "mfma_gram.hpp"
#include "wave.hpp"
#include "within.hpp"

namespace fehip {
namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 512;
constexpr int kResidBlocks = 2048;  // residual/score pass: eight workgroups per CU keep the loads in flight

using clk = std::chrono::steady_clock;

__global__ __launch_bounds__(kBlock) void wsum2_kernel(const double* __restrict__ y, const uint8_t* __restrict__ w,
                                                       int64_t n, double shift, double* part) {
    // part[b] = (sum w*(y-shift), sum w*(y-shift)^2, sum w)
    __shared__ double ws[3][kBlock / kWave];
    double a = 0.0, b = 0.0, c = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (w && !w[i]) continue;
        const double v = y[i] - shift;
        a += v;
        b += v * v;
        c += 1.0;
    }
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int wv = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        ws[0][wv] = a;
        ws[1][wv] = b;
        ws[2][wv] = c;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int k = 0; k < kBlock / kWave; ++k) t += ws[threadIdx.x][k];
        part[blockIdx.x * 3 + threadIdx.x] = t;
    }
}

// r = y~ - W~ beta on the work buffer, r'r partials per block, and (T != nullptr) the score rows
// T = r W~ (row-major n x K) in the same pass over the columns.  With K <= 16 the block's score rows
// are staged in LDS and leave as one contiguous stretch.
// cols: device array of the 1 + K column pointers [y~ W~] (panel order, any storage); the scores
// T = r x (row-major n x Kt) are formed from the Kt columns tcols (W~, or the 2SLS instrument set).
template <typename TX>
__global__ __launch_bounds__(256) void resid_scores_kernel(const TX* const* __restrict__ cols, int64_t n, int K,
                                                           const double* __restrict__ beta, double* r, double* T,
                                                           int Kt, const TX* const* __restrict__ tcols, double* part) {
    extern __shared__ double tstage[];  // [256][Kt] when staging
    const bool stage = T && Kt <= 16;
    double ss = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        const int64_t i = i0 + threadIdx.x;
        if (i < n) {
            double v = (double)cols[0][i];
            for (int k = 0; k < K; ++k) v -= (double)cols[1 + k][i] * beta[k];
            r[i] = v;
            ss += v * v;
            if (T) {
                for (int k = 0; k < Kt; ++k) {
                    const double t = v * (double)tcols[k][i];
                    if (stage)
                        tstage[threadIdx.x * Kt + k] = t;
                    else
                        T[i * Kt + k] = t;
                }
            }
        }
        if (stage) {
            __syncthreads();
            const int64_t rows = n - i0 < (int64_t)blockDim.x ? n - i0 : (int64_t)blockDim.x;
            for (int64_t e = threadIdx.x; e < rows * Kt; e += blockDim.x) T[i0 * Kt + e] = tstage[e];
            __syncthreads();
        }
    }
    __shared__ double ws[256 / 64];
    ss = wave_sum(ss);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x / 64] = ss;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 256 / 64; ++w) t += ws[w];
        part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBlock) void dot_part_kernel(const double* __restrict__ a, int64_t n, double* part) {
    __shared__ double ws[kBlock / kWave];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += a[i] * a[i];
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kBlock / kWave; ++k) t += ws[k];
        part[blockIdx.x] = t;
    }
}

// T[i][j] = r_i X[i][j] (row-major n x K) from column-major X (ld)
// T = r x, one row per thread (the columns read coalesced, each row written as K contiguous values)
template <typename TX>
__global__ void scores_rows_kernel(const double* __restrict__ r, const TX* const* __restrict__ X, int64_t n, int K,
                                   double* T) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double ri = r[i];
        for (int k = 0; k < K; ++k) T[i * K + k] = ri * (double)X[k][i];
    }
}

// S[c][j] = sum of T rows of cluster c (rows order[off[c]..off[c+1]) or identity)
__global__ void segsum_kernel(const double* __restrict__ T, int K, const int* __restrict__ order,
                              const int64_t* __restrict__ off, int64_t nS, double* S) {
    const int64_t total = nS * K;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / K;
        const int j = (int)(e - c * K);
        double s = 0.0;
        const int64_t b = off ? off[c] : c, f = off ? off[c + 1] : c + 1;
        for (int64_t k = b; k < f; ++k) s += T[(int64_t)(order ? order[k] : k) * K + j];
        S[e] = s;
    }
}

// Non-empty groups of every dim under a row subsample in one pass: flags of dim d at off[d] + id.
struct DimIds {
    const int* ids[FEHIP_MAX_FE];
    int64_t off[FEHIP_MAX_FE + 1];
    int m;
};

__global__ void nonempty_dims_kernel(DimIds dk, const uint8_t* __restrict__ w, int64_t n, int* flags) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (w[i])
            for (int d = 0; d < dk.m; ++d) flags[dk.off[d] + dk.ids[d][i]] = 1;
}

// out[d] = set flags of dim d (flags laid out as in nonempty_dims_kernel)
__global__ __launch_bounds__(kBlock) void count_dims_kernel(const int* __restrict__ f, DimIds dk,
                                                            unsigned long long* out) {
    const int64_t total = dk.off[dk.m];
    unsigned long long c[FEHIP_MAX_FE] = {0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        if (f[i])
            for (int d = 0; d < dk.m; ++d)
                if (i >= dk.off[d] && i < dk.off[d + 1]) ++c[d];
    for (int d = 0; d < dk.m; ++d) {
        unsigned long long v = c[d];
        for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
        if ((threadIdx.x & (kWave - 1)) == 0 && v) atomicAdd(out + d, v);
    }
}

__global__ __launch_bounds__(kBlock) void count_kernel(const int* __restrict__ f, int64_t n, unsigned long long* out) {
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c += f[i] ? 1 : 0;
    for (int off = kWave / 2; off > 0; off >>= 1) c += __shfl_xor(c, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(out, c);
}

// mark[key] = 1 for clusters (sorted-position keys) holding a weighted row
__global__ void mark_kernel(const int* __restrict__ keys, const int* __restrict__ order, const uint8_t* __restrict__ w,
                            int64_t n, int* mark) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        if (w[order ? order[k] : k]) mark[keys[k]] = 1;
}

__global__ void weights_panel_kernel(const uint8_t* __restrict__ w, const int* __restrict__ perm, int64_t n,
                                     uint8_t* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = w[perm ? perm[i] : i] ? 1 : 0;
}

// Upper triangle of sum_c S_c S_c' with S_c = sum of the T rows of cluster c (T row-major n x K);
// clusters are rows order[off[c]..off[c+1]) (order == nullptr: identity; off == nullptr: each row
// is its own cluster).  One thread per cluster, KT >= K accumulators per thread, block partials.
template <int KT>
__global__ __launch_bounds__(kBlock) void meat_kernel(const double* __restrict__ T, int K, int ldT,
                                                      const int* __restrict__ order, const int64_t* __restrict__ off,
                                                      int64_t nS, double* part) {
    constexpr int NP = KT * (KT + 1) / 2;
    double acc[NP];
#pragma unroll
    for (int e = 0; e < NP; ++e) acc[e] = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nS; c += (int64_t)gridDim.x * blockDim.x) {
        double sc[KT];
#pragma unroll
        for (int k = 0; k < KT; ++k) sc[k] = 0.0;
        const int64_t b = off ? off[c] : c, e = off ? off[c + 1] : c + 1;
        for (int64_t q = b; q < e; ++q) {
            const double* row = T + (int64_t)(order ? order[q] : q) * ldT;
#pragma unroll
            for (int k = 0; k < KT; ++k)
                if (k < K) sc[k] += row[k];
        }
        int idx = 0;
#pragma unroll
        for (int a = 0; a < KT; ++a)
#pragma unroll
            for (int bb = a; bb < KT; ++bb) acc[idx++] += sc[a] * sc[bb];
    }
    __shared__ double ws[kBlock / kWave][NP];
    const int wv = threadIdx.x / kWave, ln = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int e = 0; e < NP; ++e) {
        const double v = wave_sum(acc[e]);
        if (ln == 0) ws[wv][e] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NP; e += blockDim.x) {
        double v = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) v += ws[w][e];
        part[(int64_t)blockIdx.x * NP + e] = v;
    }
}

// Upper triangle of X'X for n rows of C <= KT column-major columns (ld), block partials.
template <int KT, typename TX>
__global__ __launch_bounds__(kBlock) void gram_rows_kernel(const TX* __restrict__ X, int64_t ld, int C, int64_t n,
                                                           double* part) {
    constexpr int NP = KT * (KT + 1) / 2;
    double acc[NP];
#pragma unroll
    for (int e = 0; e < NP; ++e) acc[e] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double x[KT];
#pragma unroll
        for (int k = 0; k < KT; ++k) x[k] = k < C ? (double)X[i + k * ld] : 0.0;
        int idx = 0;
#pragma unroll
        for (int a = 0; a < KT; ++a)
#pragma unroll
            for (int bb = a; bb < KT; ++bb) acc[idx++] += x[a] * x[bb];
    }
    __shared__ double ws[kBlock / kWave][NP];
    const int wv = threadIdx.x / kWave, ln = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int e = 0; e < NP; ++e) {
        const double v = wave_sum(acc[e]);
        if (ln == 0) ws[wv][e] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NP; e += blockDim.x) {
        double v = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) v += ws[w][e];
        part[(int64_t)blockIdx.x * NP + e] = v;
    }
}

// Sum the block partials of an upper triangle (KT packing) into a full symmetric K x K matrix.
int reduce_upper(const double* dpart, int nblk, int KT, int K, std::vector<double>& M) {
    const int NP = KT * (KT + 1) / 2;
    std::vector<double> h((size_t)nblk * NP);
    FEHIP_CHECK(copy_sync(h.data(), dpart, h.size() * sizeof(double), hipMemcpyDeviceToHost), "partials");
    std::vector<double> tot(NP, 0.0);
    for (int b = 0; b < nblk; ++b)
        for (int e = 0; e < NP; ++e) tot[e] += h[(size_t)b * NP + e];
    M.assign((size_t)K * K, 0.0);
    int idx = 0;
    for (int a = 0; a < KT; ++a)
        for (int b = a; b < KT; ++b, ++idx)
            if (a < K && b < K) M[(size_t)b * K + a] = M[(size_t)a * K + b] = tot[idx];
    return 0;
}

template <int KT>
struct VecK {
    double v[KT];
};

template <int KT>
struct VecPlus {
    __device__ VecK<KT> operator()(const VecK<KT>& x, const VecK<KT>& y) const {
        VecK<KT> r;
#pragma unroll
        for (int k = 0; k < KT; ++k) r.v[k] = x.v[k] + y.v[k];
        return r;
    }
};

template <int KT>
struct GatherRow {
    const double* T;
    const int* order;
    int K;
    __device__ VecK<KT> operator()(int64_t q) const {
        VecK<KT> r;
        const double* row = T + (int64_t)(order ? order[q] : q) * K;
#pragma unroll
        for (int k = 0; k < KT; ++k) r.v[k] = k < K ? row[k] : 0.0;
        return r;
    }
};

// S[c] = sum of the T rows of cluster c, for a partition given as sorted positions (order) with
// grouped keys; rocPRIM reduce-by-key handles any mix of cluster sizes (5M-row time clusters and
// 1-row intersections alike).  S is nS x KT row-major.
template <int KT>
int cluster_sums(const double* T, int K, const int* order, const int* keys, int64_t n, double* S, int64_t* d_count,
                 hipStream_t s) {
    auto vals = rocprim::make_transform_iterator(rocprim::make_counting_iterator<int64_t>(0),
                                                 GatherRow<KT>{T, order, K});
    size_t bytes = 0;
    FEHIP_CHECK(rocprim::reduce_by_key(nullptr, bytes, keys, vals, (size_t)n, rocprim::make_discard_iterator(),
                                       reinterpret_cast<VecK<KT>*>(S), d_count, VecPlus<KT>(),
                                       rocprim::equal_to<int>(), s),
                "reduce_by_key size");
    void* tmp = workspace(bytes, 8);
    if (!tmp) return set_error("reduce_by_key workspace");
    FEHIP_CHECK(rocprim::reduce_by_key(tmp, bytes, keys, vals, (size_t)n, rocprim::make_discard_iterator(),
                                       reinterpret_cast<VecK<KT>*>(S), d_count, VecPlus<KT>(),
                                       rocprim::equal_to<int>(), s),
                "reduce_by_key");
    return 0;
}

// ---- cluster meat by subset class -------------------------------------------------------------
// Subsets containing the primary dim: every cluster lies inside one primary segment, so a tile of
// whole segments (<= 256 rows, staged in LDS) holds each of its clusters entirely.  A row that is
// the first of its cluster in the segment sums the cluster's rows and adds S S' to its registers.
// One pass over the wave tiles (whole primary segments of <= 64 rows, one row per lane) forms the
// cluster sums of every subset whose clusters a tile's rows determine:
//  - subsets containing the primary dim P: a cluster never crosses a primary segment, so the scores
//    t = r x are summed per run of equal subset ids by a segmented DPP scan, runs of one segment with
//    the same ids are merged (a cluster inside a segment need not be contiguous), and each cluster's
//    leader lane adds S S' to matrix-core accumulators in registers (up to four such subsets);
//  - a single large dim other than P: the sums of t over the runs of equal id inside each segment are
//    stored at the run's place in group order (the panel's run numbering), so the groups' sums are
//    contiguous segment sums afterwards (up to two such dims);
//  - subsets of small dims only: t is added into per-block LDS tables indexed by the mixed-radix cell
//    of the row's ids, with a count of the rows in the subsample (up to three such subsets).
// The residual is read, or formed in FP64 as y~ - W~ beta from the stored columns, and the block's sum
// of r^2 is written.  The other cluster dims are loaded per row from up to three slots.
constexpr int kMaxNestSubsets = 4;
constexpr int kMaxRunOuts = 2;
constexpr int kMaxSmallOuts = 3;
constexpr int kScoreSmallLds = 4096;  // doubles of LDS for the small-cell tables of one block

struct WaveScores {
    const double* r;            // residual (panel order), or nullptr: formed from y and beta
    const void* y;              // y~ column (the columns' type) when r is null
    const double* beta;         // [K] when r is null
    double betah[12];           // the same coefficients by value: kernel arguments, read once per wave (loads
                                // through beta are repeated every tile, as the pass's stores may alias them)
    int K;
    const int* idP;
    const int64_t* wtile_rec;   // first row | rows << 32 of each wave tile
    int64_t n_wtiles;
    const int* kd[3];           // ids of the other cluster dims (nullptr: slot unused)
    int nsub;
    int use[kMaxNestSubsets];   // per subset: bit j set -> slot j belongs to the subset
    int mode[kMaxNestSubsets];  // 2: singletons, 1: clusters are single runs, 0: merge runs of a segment
    int nrun;
    int rslot[kMaxRunOuts];     // slot of the run dim
    int rshare[kMaxRunOuts];    // subset whose run scan is this dim's (its use is this slot alone), or -1
    const int* rdest[kMaxRunOuts];  // per row: the run's place in group order at run ends (-1 inside)
    double* runbuf[kMaxRunOuts];    // [runs][KT]
    int nsmall;
    int sstride[kMaxSmallOuts][3];  // per slot: the radix of its ids in the cell index (0: not in the subset)
    int soff[kMaxSmallOuts];        // LDS offset (doubles) of the subset's table [cells][KT + 1]
    int small_lds;                  // doubles of LDS tables
    double* small_part;             // [grid][small_lds]
    double* rss_part;               // [grid]
    const uint8_t* w;               // panel-order weights (nullptr: all rows)
    const int* spos;                // sorted position of each row in one subset's partition (nullptr: none)
    void* tsort;                    // the score rows in that order (row-major n x K)
    int tsort32;                    // 1: those rows in FP32 (the scores of FP32 columns, for the run-cells route)
    int* tkey;                      // with spos: 2 * (the row's cell of the small dims) + (row weighted), or nullptr
    int cstride[3];                 // per slot: the radix of its ids in that cell index (0: not in it)
};

// Row data of one wave tile as loaded: the loads are issued one tile ahead of their use, without a
// branch and from clamped addresses (rows past the tile read its first row; unused slots read the
// primary ids), and nothing is computed from them until the tile is processed, so the wait for them
// falls an iteration after their issue.  The tile records are loaded one tile earlier still.
template <int KT, typename TX>
struct ScoreRows {
    int64_t rec;
    int ip, k[3], rd[kMaxRunOuts], sp;
    uint8_t w;
    double r;  // residual when read; with FP64 columns also y~ when the residual is formed
    TX y;      // FP32 columns: y~ when the residual is formed
    TX x[KT];
};

__device__ __forceinline__ int64_t load_score_rec(const WaveScores& a, int64_t t) {
    return a.wtile_rec[t < a.n_wtiles ? t : a.n_wtiles - 1];
}

template <int KT, typename TX>
__device__ __forceinline__ void load_score_rows(const WaveScores& a, const TX* const (&xk)[KT], const TX* yc,
                                                int64_t rec, int lane, ScoreRows<KT, TX>& q) {
    q.rec = rec;
    const int64_t R0 = rec & 0xffffffffll;
    const int nr = (int)(rec >> 32);
    const int64_t ri = R0 + (lane < nr ? lane : 0);
    q.ip = a.idP[ri];
#pragma unroll
    for (int j = 0; j < 3; ++j) q.k[j] = (a.kd[j] ? a.kd[j] : a.idP)[ri];
#pragma unroll
    for (int j = 0; j < kMaxRunOuts; ++j) q.rd[j] = (j < a.nrun ? a.rdest[j] : a.idP)[ri];
    q.w = (a.w ? a.w : reinterpret_cast<const uint8_t*>(a.idP))[ri];
    q.sp = (a.spos ? a.spos : a.idP)[ri];
    if constexpr (std::is_same<TX, double>::value)
        q.r = (a.r ? a.r : yc)[ri];
    else if (a.r)
        q.r = a.r[ri];
    else
        q.y = yc[ri];
#pragma unroll
    for (int k = 0; k < KT; ++k) q.x[k] = xk[k][ri];
}

// Segmented sums of the score values over [run start rs, lane], in the scores' type: with FP64 columns an
// FP64 scan; with FP32 columns (the mixed-precision form, whose scores carry the storage rounding anyway) an
// FP32 scan of one fused DPP multiply-add per column and step.
template <int KT, typename TV>
__device__ __forceinline__ void scan_scores(TV (&sv)[KT], const TV (&tv)[KT], int lane, int rs) {
#pragma unroll
    for (int k = 0; k < KT; ++k) sv[k] = tv[k];
    if constexpr (std::is_same<TV, float>::value)
        seg_scan_cols_f32<KT>(sv, ScanFactorsF(ScanMask(lane, rs)));
    else
        seg_scan_cols<KT>(sv, ScanFactors(ScanMask(lane, rs)));
}

// Per subset s: part[(s * gridDim.x + block) * NP + e] (upper triangle of the block's sum of S S') and
// ucount[s] (non-empty clusters); run sums into runbuf; small tables into small_part; r^2 into rss_part.
// NS: compile-time bound on the nested subsets (a.nsub <= NS), which sizes their accumulators
template <int KT, typename TX, int NS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(KT <= 8 ? (NS == 1 ? 4 : 3) : 1))) void wave_scores_kernel(WaveScores a, const TX* const* __restrict__ X, double* part,
                                                             unsigned long long* ucount) {
    __shared__ double gst[(kBlock / kWave) * gram_stage<KT>()];  // per-wave staging of the cluster sums
    extern __shared__ double stab[];                            // small-cell tables [small_lds]
    __shared__ double rss_w[kBlock / kWave];
    for (int e = threadIdx.x; e < a.small_lds; e += blockDim.x) stab[e] = 0.0;
    __syncthreads();
    gram4_t acc[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[q] = gram4_t{0.0, 0.0, 0.0, 0.0};
    unsigned long long nclusters[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) nclusters[q] = 0;
    double rss = 0.0;
    const int lane = threadIdx.x & (kWave - 1);
    double* mystage = gst + (size_t)(threadIdx.x / kWave) * gram_stage<KT>();
    const int64_t wstride = (int64_t)gridDim.x * (kBlock / kWave);
    const TX* xk[KT];  // columns past K read the first one (values unused)
#pragma unroll
    for (int k = 0; k < KT; ++k) xk[k] = X[k < a.K ? k : 0];
    const TX* yc = a.r ? xk[0] : static_cast<const TX*>(a.y);
    using TV = TX;  // type of the scores and their segmented sums
    int64_t t = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    ScoreRows<KT, TX> cur, nxt;
    load_score_rows<KT, TX>(a, xk, yc, load_score_rec(a, t), lane, cur);
    int64_t rec_n = load_score_rec(a, t + wstride);
    for (; t < a.n_wtiles; t += wstride) {
        const int nr = (int)(cur.rec >> 32);
        const bool valid = lane < nr;
        const int ip = valid ? cur.ip : -1;
        int kk[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) kk[j] = (valid && a.kd[j]) ? cur.k[j] : -1;
        int rdl[kMaxRunOuts];
#pragma unroll
        for (int j = 0; j < kMaxRunOuts; ++j) rdl[j] = cur.rd[j];
        const bool wrow = valid && (!a.w || cur.w);
        const int64_t spr = cur.sp;
        // the residual in FP64 (it is the small difference of y~ and W~ beta); the scores in the columns'
        // type (FP32 storage: one rounding of r x, as the storage rounded x; the sums over clusters and the
        // cross products stay FP64)
        double rr;
        if (a.r) {
            rr = cur.r;
        } else {
            if constexpr (std::is_same<TX, double>::value)
                rr = cur.r;
            else
                rr = (double)cur.y;
#pragma unroll
            for (int k = 0; k < KT; ++k)
                if (k < a.K) rr = fma(-a.betah[k], (double)cur.x[k], rr);
        }
        rr = valid ? rr : 0.0;
        rss = fma(rr, rr, rss);
        const TV rv = (TV)rr;
        TV tv[KT];
#pragma unroll
        for (int k = 0; k < KT; ++k) tv[k] = k < a.K ? rv * cur.x[k] : (TV)0;
        // the next tiles' record and rows
        const int64_t rec_nn = load_score_rec(a, t + 2 * wstride);
        load_score_rows<KT, TX>(a, xk, yc, rec_n, lane, nxt);
        rec_n = rec_nn;
        cur = nxt;
        __builtin_amdgcn_sched_barrier(0);
        // the score row at its place in the sorted partition of the subset that reads it in that order
        bool ts32 = false;  // FP32 score rows (uniform)
        if constexpr (std::is_same<TV, float>::value) ts32 = a.tsort32 != 0;
        if constexpr (KT == 4 || KT == 8) {
            if (a.spos && a.K == KT) {
                // Full rows go through the wave's staging area, so that one store instruction writes
                // 64 / (pieces per row) whole rows, one 16-byte piece per lane: the rows of a run land next to
                // each other in the partition, and such a store covers them with whole cache lines.
                if (ts32) {
                    constexpr int P = KT / 4, RPI = kWave / P;  // pieces per row, rows per instruction
                    float4* st = reinterpret_cast<float4*>(mystage);
#pragma unroll
                    for (int k = 0; k < P; ++k)
                        st[lane * P + k] = float4{(float)tv[4 * k], (float)tv[4 * k + 1], (float)tv[4 * k + 2],
                                                  (float)tv[4 * k + 3]};
                    __builtin_amdgcn_wave_barrier();
                    float* ts = static_cast<float*>(a.tsort);
#pragma unroll
                    for (int i = 0; i < kWave / RPI; ++i) {
                        const int row = i * RPI + lane / P, piece = lane % P;
                        const int pos = __shfl((int)spr, row, kWave);
                        if (row < nr) reinterpret_cast<float4*>(ts + (int64_t)pos * KT)[piece] = st[row * P + piece];
                    }
                    __builtin_amdgcn_wave_barrier();
                } else {
                    constexpr int P = KT / 2, RPI = kWave / P;
                    double2* st = reinterpret_cast<double2*>(mystage);
#pragma unroll
                    for (int k = 0; k < P; ++k) st[lane * P + k] = double2{(double)tv[2 * k], (double)tv[2 * k + 1]};
                    __builtin_amdgcn_wave_barrier();
                    double* ts = static_cast<double*>(a.tsort);
#pragma unroll
                    for (int i = 0; i < kWave / RPI; ++i) {
                        const int row = i * RPI + lane / P, piece = lane % P;
                        const int pos = __shfl((int)spr, row, kWave);
                        if (row < nr) reinterpret_cast<double2*>(ts + (int64_t)pos * KT)[piece] = st[row * P + piece];
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
        if (a.spos && valid) {
            if ((KT == 4 || KT == 8) && a.K == KT) {
                // written above
            } else if (ts32) {
                float* dst = static_cast<float*>(a.tsort) + spr * a.K;
#pragma unroll
                for (int k = 0; k < KT; ++k)
                    if (k < a.K) dst[k] = (float)tv[k];
            } else if (a.K == KT) {  // KT is even: 16-byte stores
                double* dst = static_cast<double*>(a.tsort) + spr * a.K;
#pragma unroll
                for (int k = 0; k < KT / 2; ++k)
                    reinterpret_cast<double2*>(dst)[k] = double2{(double)tv[2 * k], (double)tv[2 * k + 1]};
            } else {
                double* dst = static_cast<double*>(a.tsort) + spr * a.K;
#pragma unroll
                for (int k = 0; k < KT; ++k)
                    if (k < a.K) dst[k] = (double)tv[k];
            }
            if (a.tkey) {
                int cell = 0;
#pragma unroll
                for (int j = 0; j < 3; ++j) cell += kk[j] * a.cstride[j];
                a.tkey[spr] = 2 * cell + (wrow ? 1 : 0);
            }
        }
        const int ipp = __shfl_up(ip, 1, kWave);
        int kp[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) kp[j] = __shfl_up(kk[j], 1, kWave);
        const bool hs = valid && (lane == 0 || ip != ipp);
        const unsigned long long H = __ballot(hs);
        const int ss = 63 - __clzll(H & mask_le(lane));  // segment start lane
        const unsigned long long BW = __ballot(wrow);
        // small-dim subsets: the row's scores into its cell
        for (int q = 0; q < a.nsmall; ++q) {
            int cell = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) cell += kk[j] * a.sstride[q][j];
            if (valid) {
                double* row = stab + a.soff[q] + cell * (KT + 1);
#pragma unroll
                for (int k = 0; k < KT; ++k)
                    if (k < a.K) lds_add(&row[k], (double)tv[k]);
                if (wrow) lds_add(&row[KT], 1.0);
            }
        }
        // run sums of a single large dim: the runs of equal id inside the segment
        auto store_runs = [&](int j, const TV (&sv)[KT], bool rend) {
            if (rend) {  // KT is even and the rows are KT doubles: 16-byte stores
                double2* dst = reinterpret_cast<double2*>(a.runbuf[j] + (int64_t)rdl[j] * KT);
#pragma unroll
                for (int k = 0; k < KT / 2; ++k) dst[k] = double2{(double)sv[2 * k], (double)sv[2 * k + 1]};
            }
        };
        for (int j = 0; j < a.nrun; ++j) {
            if (a.rshare[j] >= 0) continue;  // stored from the subset that shares its scan
            const int s0 = a.rslot[j];
            const int kj = s0 == 0 ? kk[0] : (s0 == 1 ? kk[1] : kk[2]);
            const int kjp = s0 == 0 ? kp[0] : (s0 == 1 ? kp[1] : kp[2]);
            const bool rh = valid && (hs || kj != kjp);
            const unsigned long long HR = __ballot(rh);
            const int rs = 63 - __clzll(HR & mask_le(lane));
            const bool rend = valid && (lane == nr - 1 || (lane < kWave - 1 && ((HR >> (lane + 1)) & 1ull)));
            TV sv[KT];
            scan_scores<KT, TV>(sv, tv, lane, rs);
            store_runs(j, sv, rend);
        }
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            if (q >= a.nsub) break;
            const int use = a.use[q], mode = a.mode[q];
            int ks[3], ksp[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                ks[j] = ((use >> j) & 1) ? kk[j] : -1;
                ksp[j] = ((use >> j) & 1) ? kp[j] : -1;
            }
            const bool rh = valid && (hs || ks[0] != ksp[0] || ks[1] != ksp[1] || ks[2] != ksp[2]);
            const unsigned long long HR = __ballot(rh);
            const unsigned long long le = HR & mask_le(lane);
            const int rs = 63 - __clzll(le);
            const bool rend = valid && (lane == nr - 1 || (lane < kWave - 1 && ((HR >> (lane + 1)) & 1ull)));
            bool ne = (BW & mask_le(lane) & ~((1ull << rs) - 1ull)) != 0;  // run holds a weighted row
            TV sv[KT];
            if (mode != 2) {
                scan_scores<KT, TV>(sv, tv, lane, rs);
            } else {
#pragma unroll
                for (int k = 0; k < KT; ++k) sv[k] = tv[k];
            }
            for (int j = 0; j < a.nrun; ++j)
                if (a.rshare[j] == q) store_runs(j, sv, rend);
            bool leader = rend;
            if (use != 0 && mode == 0) {
                // merge the runs of one segment that carry the same ids
                unsigned long long E = __ballot(rend);
                while (E) {
                    const int j = __ffsll((long long)E) - 1;
                    E &= E - 1ull;
                    const int kj0 = __shfl(ks[0], j, kWave), kj1 = __shfl(ks[1], j, kWave),
                              kj2 = __shfl(ks[2], j, kWave), sj = __shfl(ss, j, kWave);
                    const bool match =
                        rend && lane != j && sj == ss && kj0 == ks[0] && kj1 == ks[1] && kj2 == ks[2];
                    if (match && j < lane) leader = false;
                    const bool take = match && j > lane;
                    if (__ballot(take)) {
#pragma unroll
                        for (int k = 0; k < KT; ++k) {
                            const TV vj = __shfl(sv[k], j, kWave);
                            if (take) sv[k] += vj;
                        }
                        const bool nej = __shfl((int)ne, j, kWave) != 0;
                        if (take) ne = ne || nej;
                    }
                }
            }
            const bool on = leader && ne;
            const unsigned long long BU = __ballot(on);
            if (lane == 0) nclusters[q] += (unsigned long long)__popcll(BU);
            // sum of S S' over the clusters that end here: the cross products of their sums, packed (a subset
            // of singletons keeps every row)
            gram_wave_sparse<KT>(mystage, sv, on, lane, acc[q]);
        }
    }
    constexpr int NP = KT * (KT + 1) / 2;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        if (q >= a.nsub) break;
        if (lane == 0 && nclusters[q]) atomicAdd(ucount + q, nclusters[q]);
        gram_store_upper<KT>(gst, gram_stage<KT>(), acc[q], part + ((int64_t)q * gridDim.x + blockIdx.x) * NP);
    }
    rss = wave_sum(rss);
    if (lane == 0) rss_w[threadIdx.x / kWave] = rss;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t2 = 0.0;
        for (int v = 0; v < kBlock / kWave; ++v) t2 += rss_w[v];
        a.rss_part[blockIdx.x] = t2;
    }
    for (int e = threadIdx.x; e < a.small_lds; e += blockDim.x)
        a.small_part[(int64_t)blockIdx.x * a.small_lds + e] = stab[e];
}

// sums of the run partials of each group (contiguous in group order): S[g][c] for c < KT, four loads in
// flight per thread
template <int KT>
__global__ void group_sum_kernel(const double* __restrict__ runbuf, const int64_t* __restrict__ goff, int G,
                                 double* S) {
    const int64_t total = (int64_t)G * KT;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int g = (int)(e / KT), c = (int)(e % KT);
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        const int64_t k0 = goff[g], k1 = goff[g + 1];
        int64_t k = k0;
        for (; k + 4 <= k1; k += 4) {
            s0 += runbuf[k * KT + c];
            s1 += runbuf[(k + 1) * KT + c];
            s2 += runbuf[(k + 2) * KT + c];
            s3 += runbuf[(k + 3) * KT + c];
        }
        for (; k < k1; ++k) s0 += runbuf[k * KT + c];
        S[e] = (s0 + s1) + (s2 + s3);
    }
}

// Any subset through its cached sorted partition, one wave per 64 consecutive sorted positions: the
// score rows of the positions are summed per cluster by a segmented DPP scan; a cluster that starts
// and ends inside the wave adds S S' to its end lane's registers.  A cluster cut by a wave boundary
// leaves its pieces in two boundary slots per wave (slot 0: the piece before the wave's first cluster
// head, or the whole wave when it has no head; slot 1: the last cluster when it runs on into the next
// wave), keyed by cluster id; the pieces are merged by a reduce-by-key afterwards.  An empty slot
// has zero sums and key -1, except that slot 1 of a headless wave repeats slot 0's key so that a
// cluster's pieces stay adjacent.
template <int KT>
__global__ __launch_bounds__(kBlock) void sorted_wave_meat_kernel(const double* __restrict__ T, int K,
                                                                  const int* __restrict__ order,
                                                                  const int* __restrict__ keys, int64_t n,
                                                                  double* part, int* bkey, double* bval) {
    __shared__ double gst[(kBlock / kWave) * gram_stage<KT>()];  // per-wave staging of the cluster sums
    gram4_t acc = {0.0, 0.0, 0.0, 0.0};
    const int lane = threadIdx.x & (kWave - 1);
    double* mystage = gst + (size_t)(threadIdx.x / kWave) * gram_stage<KT>();
    const int64_t nwaves = (n + kWave - 1) / kWave;
    const int64_t wstride = (int64_t)gridDim.x * (kBlock / kWave);
    // Two-stage software pipeline: the keys and row indices of the tile after next and the score rows of
    // the next tile are requested before the current tile is reduced, so a wave keeps its random row
    // gathers in flight while it scans.
    auto load_ids = [&](int64_t wt, int& key, int64_t& row) {
        const int64_t qq = wt * kWave + lane;
        const bool v = wt < nwaves && qq < n;
        key = v ? keys[qq] : -1;
        row = v ? (order ? (int64_t)order[qq] : qq) : -1;
    };
    auto load_rows = [&](int64_t row, double (&t)[KT]) {
        if (K == KT) {  // full rows: 16-byte loads (KT is even, so every row starts 16-byte aligned)
            const double2* src = reinterpret_cast<const double2*>(T + (row >= 0 ? row : 0) * KT);
#pragma unroll
            for (int k = 0; k < KT / 2; ++k) {
                const double2 v = row >= 0 ? src[k] : double2{0.0, 0.0};
                t[2 * k] = v.x;
                t[2 * k + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int k = 0; k < KT; ++k) t[k] = (row >= 0 && k < K) ? T[row * K + k] : 0.0;
        }
    };
    int64_t w = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    int key_c, key_n;
    int64_t row_c, row_n;
    double sv_n[KT];
    load_ids(w, key_c, row_c);
    load_ids(w + wstride, key_n, row_n);
    load_rows(row_c, sv_n);
    for (; w < nwaves; w += wstride) {
        double sv[KT];
#pragma unroll
        for (int k = 0; k < KT; ++k) sv[k] = sv_n[k];
        const int key0 = key_c;
        key_c = key_n;
        row_c = row_n;
        load_rows(row_c, sv_n);                   // score rows of the next tile
        load_ids(w + 2 * wstride, key_n, row_n);  // ids of the tile after next
        const int64_t q0 = w * kWave;
        const int nvalid = (int)(n - q0 < kWave ? n - q0 : kWave);
        const bool valid = lane < nvalid;
        const int key = valid ? key0 : -1;
        int kprev = __shfl_up(key, 1, kWave);
        if (lane == 0) kprev = q0 > 0 ? keys[q0 - 1] : -2;
        const bool head = valid && key != kprev;
        const unsigned long long H = __ballot(head);
        const int last = nvalid - 1;
        // does the cluster at the last position continue into the next wave?
        const int klast = __shfl(key, last, kWave);
        const bool cont = (q0 + nvalid < n) && keys[q0 + nvalid] == klast;
        const unsigned long long le = H & mask_le(lane);
        const int ss = le ? 63 - __clzll(le) : 0;
        seg_scan_cols<KT>(sv, ScanFactors(ScanMask(lane, ss)));
        const bool nexthead = lane < last ? ((H >> (lane + 1)) & 1ull) != 0 : !cont;
        const bool cend = valid && (lane <= last) && nexthead;
        const bool complete = cend && le != 0;  // started at a head of this wave
        gram_wave_sparse<KT>(mystage, sv, complete, lane, acc);  // the clusters that end here, packed
        // boundary pieces
        const bool open_left = (H & 1ull) == 0ull;
        const int e0 = H ? __ffsll((long long)H) - 2 : last;  // last lane of the leading piece
        const bool open_right = cont && H != 0ull;             // trailing cluster started in this wave
        const int64_t b0 = 2 * w, b1 = 2 * w + 1;
        if (lane == (open_left ? e0 : 0)) {
            bkey[b0] = open_left ? key : -1;
#pragma unroll
            for (int k = 0; k < KT; ++k) bval[b0 * KT + k] = open_left ? sv[k] : 0.0;
        }
        if (lane == (open_right ? last : 0)) {
            bkey[b1] = open_right ? key : (H == 0ull ? klast : -1);
#pragma unroll
            for (int k = 0; k < KT; ++k) bval[b1 * KT + k] = open_right ? sv[k] : 0.0;
        }
    }
    gram_store_upper<KT>(gst, gram_stage<KT>(), acc, part + (int64_t)blockIdx.x * (KT * (KT + 1) / 2));
}

// A subset of one large dim A (not the primary) and small dims B with at most kMaxRunCells cells.  The
// fused pass writes each row's scores, and its cell of B (key = 2 * cell + weighted), at the row's place
// in A's partition: the partition sorts stably by A, so the groups of A are contiguous and the rows of a
// run of A inside a primary segment stay together, and the pass writes whole runs instead of single rows
// to random places.  One wave per group of A adds the group's rows into an LDS table with a row per cell
// (plus a count of weighted rows), then adds S S' of the cells that hold a weighted row to its
// matrix-core accumulator and clears the table.
constexpr int kMaxRunCells = 64;

template <int KT, typename TT>
__global__ __launch_bounds__(kBlock) void run_cells_meat_kernel(const TT* __restrict__ T, int K,
                                                                const int* __restrict__ tkey,
                                                                const int64_t* __restrict__ off, int64_t G,
                                                                int ncells, double* part,
                                                                unsigned long long* n_clusters) {
    __shared__ double gst[(kBlock / kWave) * gram_stage<KT>()];
    __shared__ double tab[(kBlock / kWave) * kMaxRunCells * (KT + 1)];
    gram4_t acc = {0.0, 0.0, 0.0, 0.0};
    unsigned long long ncl = 0;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    double* mystage = gst + (size_t)wv * gram_stage<KT>();
    double* mytab = tab + (size_t)wv * kMaxRunCells * (KT + 1);
    for (int e = lane; e < kMaxRunCells * (KT + 1); e += kWave) mytab[e] = 0.0;
    const int64_t wstride = (int64_t)gridDim.x * (kBlock / kWave);
    for (int64_t g = (int64_t)blockIdx.x * (kBlock / kWave) + wv; g < G; g += wstride) {
        const int64_t a = off[g], b = off[g + 1];
        __builtin_amdgcn_wave_barrier();
        for (int64_t q = a; q < b; q += kWave) {
            const int64_t i = q + lane;
            if (i < b) {
                const int key = tkey[i];
                double* row = mytab + min(max(key >> 1, 0), ncells - 1) * (KT + 1);
                double v[KT];
                if (K == KT) {  // rows start 16-byte aligned: FP64 pairs, or FP32 quadruples (KT is 4, 8 or 12)
                    if constexpr (sizeof(TT) == 4) {
                        const float4* src = reinterpret_cast<const float4*>(T + i * KT);
#pragma unroll
                        for (int k = 0; k < KT / 4; ++k) {
                            const float4 x = src[k];
                            v[4 * k] = x.x;
                            v[4 * k + 1] = x.y;
                            v[4 * k + 2] = x.z;
                            v[4 * k + 3] = x.w;
                        }
                    } else {
                        const double2* src = reinterpret_cast<const double2*>(T + i * KT);
#pragma unroll
                        for (int k = 0; k < KT / 2; ++k) {
                            const double2 x = src[k];
                            v[2 * k] = x.x;
                            v[2 * k + 1] = x.y;
                        }
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < KT; ++k) v[k] = k < K ? (double)T[i * K + k] : 0.0;
                }
#pragma unroll
                for (int k = 0; k < KT; ++k)
                    if (k < K) lds_add(&row[k], v[k]);
                if (key & 1) lds_add(&row[KT], 1.0);
            }
        }
        __builtin_amdgcn_wave_barrier();
        double sv[KT];
        bool on = false;
        if (lane < ncells) {
            double* row = mytab + lane * (KT + 1);
            on = row[KT] > 0.0;
#pragma unroll
            for (int k = 0; k < KT; ++k) sv[k] = on ? row[k] : 0.0;
#pragma unroll
            for (int k = 0; k <= KT; ++k) row[k] = 0.0;
        } else {
#pragma unroll
            for (int k = 0; k < KT; ++k) sv[k] = 0.0;
        }
        const unsigned long long BU = __ballot(on);
        if (lane == 0) ncl += (unsigned long long)__popcll(BU);
        gram_wave_sparse<KT>(mystage, sv, on, lane, acc);  // the cells with a weighted row, packed
    }
    if (lane == 0 && ncl) atomicAdd(n_clusters, ncl);
    gram_store_upper<KT>(gst, gram_stage<KT>(), acc, part + (int64_t)blockIdx.x * (KT * (KT + 1) / 2));
}

struct NestKeys {
    const int* ids[3];  // other dims of the subset (panel order)
    int nk;
};

template <int KT>
__global__ __launch_bounds__(kBlock) void nested_meat_kernel(const double* __restrict__ T, int K,
                                                             const int64_t* __restrict__ seg_off,
                                                             const int64_t* __restrict__ tile_seg, int64_t n_tiles,
                                                             NestKeys nk, const uint8_t* __restrict__ w, double* part,
                                                             unsigned long long* n_clusters) {
    unsigned long long nheads = 0;
    constexpr int NP = KT * (KT + 1) / 2;
    __shared__ double ts[kTileRows * KT];
    __shared__ int key[3][kTileRows];
    __shared__ int soff[kTileRows + 1];
    __shared__ unsigned char wt[kTileRows];
    double acc[NP];
#pragma unroll
    for (int e = 0; e < NP; ++e) acc[e] = 0.0;
    const int j = threadIdx.x;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const int64_t s0 = tile_seg[2 * t], s1 = tile_seg[2 * t + 1];
        const int nseg = (int)(s1 - s0);
        const int64_t R0 = seg_off[s0];
        __syncthreads();
        for (int k = j; k <= nseg; k += blockDim.x) soff[k] = (int)(seg_off[s0 + k] - R0);
        const int nr = (int)(seg_off[s1] - R0);
        for (int e = j; e < nr * K; e += blockDim.x) {  // coalesced copy of the tile's T rows
            const int q = e / K, c = e - q * K;
            ts[q * KT + c] = T[R0 * K + e];
        }
        if (j < nr) {
            for (int d = 0; d < nk.nk; ++d) key[d][j] = nk.ids[d][R0 + j];
            wt[j] = w ? w[R0 + j] : 1;
        }
        __syncthreads();
        if (j < nr && wt[j]) {
            int lo = 0, hi = nseg - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (soff[mid] <= j)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int b = soff[lo], f = soff[lo + 1];
            auto same = [&](int q) {
                for (int d = 0; d < nk.nk; ++d)
                    if (key[d][q] != key[d][j]) return false;
                return true;
            };
            bool first = true;
            for (int q = b; q < j && first; ++q) first = !(wt[q] && same(q));
            if (first) {
                ++nheads;
                double sc[KT];
#pragma unroll
                for (int c = 0; c < KT; ++c) sc[c] = 0.0;
                for (int q = j; q < f; ++q)
                    if (same(q))
#pragma unroll
                        for (int c = 0; c < KT; ++c) sc[c] += c < K ? ts[q * KT + c] : 0.0;
                int idx = 0;
#pragma unroll
                for (int a = 0; a < KT; ++a)
#pragma unroll
                    for (int bb = a; bb < KT; ++bb) acc[idx++] += sc[a] * sc[bb];
            }
        }
    }
    __shared__ double ws[kBlock / kWave][NP];
    const int wv = j / kWave, ln = j & (kWave - 1);
#pragma unroll
    for (int e = 0; e < NP; ++e) {
        const double v = wave_sum(acc[e]);
        if (ln == 0) ws[wv][e] = v;
    }
    __syncthreads();
    for (int e = j; e < NP; e += blockDim.x) {
        double v = 0.0;
        for (int q = 0; q < kBlock / kWave; ++q) v += ws[q][e];
        part[(int64_t)blockIdx.x * NP + e] = v;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) nheads += __shfl_xor(nheads, off, kWave);
    if (ln == 0) atomicAdd(n_clusters, nheads);
}

// Subsets of small dims only: cluster = mixed-radix index of the small ids (<= kSmallCells cells);
// per-workgroup LDS tables of T sums, written as block partials.
constexpr int kSmallCells = 1024;
struct SmallKeys {
    const int* ids[3];
    int stride[3];
    int nk;
    const uint8_t* w;
};

// Scores are read from T (row-major n x K) or, when T is null, formed as r x from the residual and the
// regressor columns (no score matrix written or read back).
template <int KT, typename TX>
__global__ __launch_bounds__(kBlock) void small_scores_kernel(const double* __restrict__ T, const double* __restrict__ r,
                                                              const TX* const* __restrict__ X, int K, int64_t n,
                                                              SmallKeys sk, int cells, double* part) {
    extern __shared__ double tab[];  // [cells][KT]
    for (int e = threadIdx.x; e < cells * KT; e += blockDim.x) tab[e] = 0.0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int cell = 0;
        for (int d = 0; d < sk.nk; ++d) cell += sk.ids[d][i] * sk.stride[d];
        if (T) {
            for (int c = 0; c < K; ++c) lds_add(&tab[cell * KT + c], T[i * K + c]);
        } else {
            const double ri = r[i];
            for (int c = 0; c < K; ++c) lds_add(&tab[cell * KT + c], ri * (double)X[c][i]);
        }
        if (!sk.w || sk.w[i]) lds_add(&tab[cell * KT + (KT - 1)], 1.0);  // occupancy (KT > K)
    }
    __syncthreads();
    for (int e = threadIdx.x; e < cells * KT; e += blockDim.x) part[(int64_t)blockIdx.x * cells * KT + e] = tab[e];
}

__global__ __launch_bounds__(kBlock) void sum_blocks_kernel(const double* __restrict__ part, int nblk, int width,
                                                            double* out) {
    __shared__ double ws[kBlock / kWave];
    const int e = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(int64_t)b * width + e];
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += ws[w];
        out[e] = t;
    }
}

// One large dim (not the primary): sums of T over the runs of equal id inside each primary tile,
// written contiguously per tile (the same run numbering the panel built), then reduced in group order.
template <int KT, typename TX>
__global__ __launch_bounds__(kBlock) void run_scores_kernel(const double* __restrict__ T, const double* __restrict__ r,
                                                            const TX* const* __restrict__ X, int K,
                                                            const int64_t* __restrict__ seg_off,
                                                            const int64_t* __restrict__ tile_seg, int64_t n_tiles,
                                                            const int* __restrict__ ids,
                                                            const int64_t* __restrict__ seg_run_off, double* runbuf) {
    __shared__ double ts[kTileRows * KT];
    __shared__ int soff[kTileRows + 1];
    __shared__ int rstart[kTileRows + 1];
    __shared__ int wcnt[kBlock / kWave];
    const int j = threadIdx.x, lane = j & (kWave - 1), wv = j / kWave;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const int64_t s0 = tile_seg[2 * t], s1 = tile_seg[2 * t + 1];
        const int nseg = (int)(s1 - s0);
        const int64_t R0 = seg_off[s0];
        __syncthreads();
        for (int k = j; k <= nseg; k += blockDim.x) soff[k] = (int)(seg_off[s0 + k] - R0);
        const int nr = (int)(seg_off[s1] - R0);
        if (T) {
            for (int e = j; e < nr * K; e += blockDim.x) {
                const int q = e / K, c = e - q * K;
                ts[q * KT + c] = T[R0 * K + e];
            }
        } else {  // column by column, coalesced
            for (int e = j; e < nr * K; e += blockDim.x) {
                const int c = e / nr, q = e - c * nr;
                ts[q * KT + c] = r[R0 + q] * (double)X[c][R0 + q];
            }
        }
        __syncthreads();
        const bool valid = j < nr;
        const int g = valid ? ids[R0 + j] : -1;
        int gp = __shfl_up(g, 1, kWave);
        if (lane == 0) gp = (valid && j > 0) ? ids[R0 + j - 1] : -1;
        bool segstart = false;
        if (valid) {
            int lo = 0, hi = nseg - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (soff[mid] <= j)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            segstart = soff[lo] == j;
        }
        const bool head = valid && (j == 0 || segstart || g != gp);
        const unsigned long long bal = __ballot(head);
        if (lane == 0) wcnt[wv] = __popcll(bal);
        __syncthreads();
        int base = 0, total = 0;
        for (int q = 0; q < kBlock / kWave; ++q) {
            base += q < wv ? wcnt[q] : 0;
            total += wcnt[q];
        }
        if (head) rstart[base + __popcll(bal & ((1ull << lane) - 1ull))] = j;
        if (j == 0) rstart[total] = nr;
        __syncthreads();
        double* dst = runbuf + seg_run_off[s0] * KT;
        for (int e = j; e < total * KT; e += blockDim.x) {
            const int k = e / KT, c = e - k * KT;
            double sm = 0.0;
            if (c < K)
                for (int q = rstart[k]; q < rstart[k + 1]; ++q) sm += ts[q * KT + c];
            dst[e] = sm;
        }
    }
}

template <int KT>
__global__ void group_runs_kernel(const double* __restrict__ runbuf, const int* __restrict__ order,
                                  const int64_t* __restrict__ goff, int G, double* S) {
    const int64_t total = (int64_t)G * KT;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int g = (int)(e / KT), c = (int)(e % KT);
        double sm = 0.0;
        for (int64_t k = goff[g]; k < goff[g + 1]; ++k) sm += runbuf[(int64_t)order[k] * KT + c];
        S[e] = sm;
    }
}

constexpr int kWideCols = 16;  // score columns per step of the wide pass: one 128-byte line of a sum row

// Scores wider than the register-tiled kernels hold (2SLS instrument sets, wide designs): one pass over
// the columns of x, in tiles of whole primary segments, forms the sums of r x over each primary segment
// (rows of Sseg, in segment order) and over each run of the large dim D (rows of Rg, at the run's
// position in D's group order), both row-major with leading dimension ld (a multiple of kWideCols;
// the columns past K are written as zeros).  A tile's rows stay in LDS for kWideCols columns at a
// time, so every column is read once and coalesced and every sum row leaves in whole 128-byte lines.
// cont[position] = 1 flags a run whose (segment, group) cell an earlier run of the same segment
// opened; that earlier run directly precedes it in group order, since group order sorts the runs of
// a group by their position in the panel.  counts[0] += cells holding a weighted row (every cell when
// w is null), counts[1] += flagged runs.
template <typename TX>
__global__ __launch_bounds__(kBlock) void wide_scores_kernel(
    const double* __restrict__ r, const TX* const* __restrict__ X, int K, int ld, const int64_t* __restrict__ seg_off,
    const int64_t* __restrict__ tile_seg, int64_t n_tiles, const int* __restrict__ ids,
    const int64_t* __restrict__ seg_run_off, const int* __restrict__ run_dest, const uint8_t* __restrict__ w,
    double* Sseg, double* Rg, uint8_t* cont, unsigned long long* counts) {
    __shared__ double ts[kWideCols][kTileRows + 1];
    __shared__ int soff[kTileRows + 1];
    __shared__ int rstart[kTileRows + 1];
    __shared__ int rdst[kTileRows];
    __shared__ int rseg[kTileRows];
    __shared__ int rgid[kTileRows];
    __shared__ unsigned char rw[kTileRows];
    __shared__ int wcnt[kBlock / kWave];
    const int j = threadIdx.x, lane = j & (kWave - 1), wv = j / kWave;
    unsigned long long ncell = 0, ncont = 0;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const int64_t s0 = tile_seg[2 * t], s1 = tile_seg[2 * t + 1];
        const int nseg = (int)(s1 - s0);
        const int64_t R0 = seg_off[s0];
        __syncthreads();  // the previous tile's reads of the shared arrays are done
        for (int k = j; k <= nseg; k += blockDim.x) soff[k] = (int)(seg_off[s0 + k] - R0);
        const int nr = (int)(seg_off[s1] - R0);
        __syncthreads();
        const bool valid = j < nr;
        const double rj = valid ? r[R0 + j] : 0.0;
        int nrun = 0;
        if (Rg) {
            // runs: a row opens one at its segment's start or where D's id changes
            const int g = valid ? ids[R0 + j] : -1;
            int seg = 0;
            if (valid) {  // the last segment starting at or before row j (empty segments start there too)
                int lo = 0, hi = nseg - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (soff[mid] <= j)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                seg = lo;
            }
            int gp = __shfl_up(g, 1, kWave);
            if (lane == 0) gp = (valid && j > 0) ? ids[R0 + j - 1] : -1;
            const bool head = valid && (soff[seg] == j || g != gp);
            const unsigned long long bal = __ballot(head);
            if (lane == 0) wcnt[wv] = __popcll(bal);
            __syncthreads();
            int base = 0;
            for (int q = 0; q < kBlock / kWave; ++q) {
                base += q < wv ? wcnt[q] : 0;
                nrun += wcnt[q];
            }
            if (head) {
                const int q = base + __popcll(bal & ((1ull << lane) - 1ull));
                rstart[q] = j;
                rseg[q] = seg;
                rgid[q] = g;
                rdst[q] = run_dest[seg_run_off[s0] + q];
            }
            if (j == 0) rstart[nrun] = nr;
            __syncthreads();
            if (j < nrun) {
                bool wq = w == nullptr;
                for (int i = rstart[j]; i < rstart[j + 1] && !wq; ++i) wq = w[R0 + i] != 0;
                rw[j] = wq ? 1 : 0;
            }
            __syncthreads();
            if (j < nrun) {
                bool c = false;
                for (int q2 = j - 1; q2 >= 0 && rseg[q2] == rseg[j] && !c; --q2) c = rgid[q2] == rgid[j];
                cont[rdst[j]] = c ? 1 : 0;
                if (c) {
                    ++ncont;
                } else {
                    bool cw = rw[j] != 0;
                    for (int q2 = j + 1; q2 < nrun && rseg[q2] == rseg[j]; ++q2)
                        if (rgid[q2] == rgid[j]) cw |= rw[q2] != 0;
                    ncell += cw ? 1 : 0;
                }
            }
        }
        // the next block of columns is loaded into registers while the current one is summed
        double nv[kWideCols];
        auto load_cols = [&](int k0) {
            const int nc = min(kWideCols, K - k0);
#pragma unroll
            for (int c = 0; c < kWideCols; ++c) nv[c] = (valid && c < nc) ? rj * (double)X[k0 + c][R0 + j] : 0.0;
        };
        load_cols(0);
        for (int k0 = 0; k0 < ld; k0 += kWideCols) {
            const int nc = min(kWideCols, K - k0);
#pragma unroll
            for (int c = 0; c < kWideCols; ++c) ts[c][j] = nv[c];
            __syncthreads();
            if (k0 + kWideCols < ld) load_cols(k0 + kWideCols);
            if (Sseg)
                for (int e = j; e < nseg * kWideCols; e += blockDim.x) {
                    const int s = e / kWideCols, c = e % kWideCols;
                    double sm = 0.0;
                    if (c < nc)
                        for (int i = soff[s]; i < soff[s + 1]; ++i) sm += ts[c][i];
                    Sseg[(s0 + s) * ld + k0 + c] = sm;
                }
            if (Rg)
                for (int e = j; e < nrun * kWideCols; e += blockDim.x) {
                    const int q = e / kWideCols, c = e % kWideCols;
                    double sm = 0.0;
                    if (c < nc)
                        for (int i = rstart[q]; i < rstart[q + 1]; ++i) sm += ts[c][i];
                    Rg[(int64_t)rdst[q] * ld + k0 + c] = sm;
                }
            __syncthreads();
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        ncell += __shfl_down(ncell, o, kWave);
        ncont += __shfl_down(ncont, o, kWave);
    }
    if (lane == 0 && Rg) {
        if (ncell) atomicAdd(counts, ncell);
        if (ncont) atomicAdd(counts + 1, ncont);
    }
}

// S[g][c] = sum of rows off[g] .. off[g+1) of R (row-major, leading dimension ld), c < ld
__global__ void row_group_sums_kernel(const double* __restrict__ R, int ld, const int64_t* __restrict__ off, int64_t G,
                                      double* S) {
    const int64_t total = G * ld;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = e / ld;
        const int c = (int)(e - g * ld);
        double sm = 0.0;
        for (int64_t k = off[g]; k < off[g + 1]; ++k) sm += R[k * ld + c];
        S[e] = sm;
    }
}

// Rows of R flagged in cont are added into the nearest unflagged row before them (their cell's first
// run) and cleared, so R then holds one sum row per cell and zero rows.
__global__ void merge_cells_kernel(double* R, int ld, const uint8_t* __restrict__ cont, int64_t nR) {
    const int64_t total = nR * ld;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = e / ld;
        if (cont[q] || q + 1 >= nR || !cont[q + 1]) continue;
        const int c = (int)(e - q * ld);
        double sm = R[e];
        for (int64_t q2 = q + 1; q2 < nR && cont[q2]; ++q2) {
            sm += R[q2 * ld + c];
            R[q2 * ld + c] = 0.0;
        }
        R[e] = sm;
    }
}

constexpr int kMeatBlocks = 1024;

// meat of one partition (K <= 12 fused path)
int meat_fused(const double* T, int K, int ldT, const int* order, const int64_t* off, int64_t nS, double* part,
               std::vector<double>& M) {
    const int KT = K <= 4 ? 4 : (K <= 8 ? 8 : 12);
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(kMeatBlocks, (nS + kBlock - 1) / kBlock));
    if (KT == 4)
        hipLaunchKernelGGL(meat_kernel<4>, dim3(grid), dim3(kBlock), 0, 0, T, K, ldT, order, off, nS, part);
    else if (KT == 8)
        hipLaunchKernelGGL(meat_kernel<8>, dim3(grid), dim3(kBlock), 0, 0, T, K, ldT, order, off, nS, part);
    else
        hipLaunchKernelGGL(meat_kernel<12>, dim3(grid), dim3(kBlock), 0, 0, T, K, ldT, order, off, nS, part);
    FEHIP_LAUNCH_CHECK("meat kernel");
    return reduce_upper(part, (int)grid, KT, K, M);
}

// Gram of C <= 12 columns (fused path)
template <typename TX>
int gram_fused(const TX* X, int64_t ld, int C, int64_t n, double* part, std::vector<double>& G) {
    const int KT = C <= 4 ? 4 : (C <= 8 ? 8 : 12);
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(kMeatBlocks, (n + kBlock - 1) / kBlock));
    if (KT == 4)
        hipLaunchKernelGGL((gram_rows_kernel<4, TX>), dim3(grid), dim3(kBlock), 0, 0, X, ld, C, n, part);
    else if (KT == 8)
        hipLaunchKernelGGL((gram_rows_kernel<8, TX>), dim3(grid), dim3(kBlock), 0, 0, X, ld, C, n, part);
    else
        hipLaunchKernelGGL((gram_rows_kernel<12, TX>), dim3(grid), dim3(kBlock), 0, 0, X, ld, C, n, part);
    FEHIP_LAUNCH_CHECK("gram kernel");
    return reduce_upper(part, (int)grid, KT, C, G);
}

// Gram of up to 16 columns on the matrix cores (the per-thread accumulators of gram_rows_kernel would
// need 136 register pairs at 16 columns): one wave per 64 consecutive rows, one block partial (upper
// triangle) per workgroup.
template <typename TX>
__global__ __launch_bounds__(kBlock) void gram_cols_kernel(const TX* const* __restrict__ cols, int C, int64_t n,
                                                           double* part) {
    __shared__ double gst[(kBlock / kWave) * gram_stage<16>()];
    gram4_t acc = {0.0, 0.0, 0.0, 0.0};
    const int lane = threadIdx.x & (kWave - 1);
    double* mystage = gst + (size_t)(threadIdx.x / kWave) * gram_stage<16>();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {  // uniform per block
        const int64_t i = i0 + threadIdx.x;
        double v[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) v[c] = (c < C && i < n) ? (double)cols[c][i] : 0.0;
        gram_wave<16>(mystage, v, lane, acc);
    }
    gram_store_upper<16>(gst, gram_stage<16>(), acc, part + (int64_t)blockIdx.x * (16 * 17 / 2));
}

// ---- FP32 storage with FP64 refinement -------------------------------------------------------
// e = y - W beta (caller order, FP64 columns)
__global__ void combo_residual_kernel(const double* __restrict__ y, const double* __restrict__ W, int64_t ld, int K,
                                      const double* __restrict__ beta, int64_t n, double* e) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = y[i];
        for (int k = 0; k < K; ++k) v -= beta[k] * W[i + (int64_t)k * ld];
        e[i] = v;
    }
}

// tables of e: a_e = a_y - sum_k beta_k a_k  (tables [C][S], column 0 = y)
__global__ void combo_tables_kernel(const double* __restrict__ alpha, int64_t S, int K, const double* __restrict__ beta,
                                    double* out) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < S; g += (int64_t)gridDim.x * blockDim.x) {
        double v = alpha[g];
        for (int k = 0; k < K; ++k) v -= beta[k] * alpha[(int64_t)(1 + k) * S + g];
        out[g] = v;
    }
}

// part[b][j] = sum_i X_j[i] e[i] for the FP32 work columns j = 1..C-1 (chunk of at most 16)
template <typename TX>
__global__ __launch_bounds__(kBlock) void xdot_kernel(const TX* __restrict__ X, int64_t n, int j0, int nj,
                                                      const double* __restrict__ e, double* part) {
    double acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double ev = e[i];
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (k < nj) acc[k] += (double)X[i + (int64_t)(j0 + k) * n] * ev;
    }
    __shared__ double ws[kBlock / kWave][16];
    const int wv = threadIdx.x / kWave, ln = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const double v = wave_sum(acc[k]);
        if (ln == 0) ws[wv][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        double v = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) v += ws[w][threadIdx.x];
        part[(int64_t)blockIdx.x * 16 + threadIdx.x] = v;
    }
}

// r = e - sum_k X_{1+k} dbeta_k (FP32 work columns)
template <typename TX>
__global__ void resid_update_kernel(const double* __restrict__ e, const TX* __restrict__ X, int64_t n, int K,
                                    const double* __restrict__ dbeta, double* r) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = e[i];
        for (int k = 0; k < K; ++k) v -= (double)X[i + (int64_t)(1 + k) * n] * dbeta[k];
        r[i] = v;
    }
}

__global__ void widen_kernel(const float* __restrict__ x, int64_t count, double* y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (double)x[i];
}

// heads of the runs of a composite key inside primary segments
__global__ __launch_bounds__(kBlock) void composite_runs_kernel(const int* __restrict__ idP, const int* __restrict__ k0,
                                                                const int* __restrict__ k1, int64_t n,
                                                                unsigned long long* count) {
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c += (i == 0 || idP[i] != idP[i - 1] || (k0 && k0[i] != k0[i - 1]) || (k1 && k1[i] != k1[i - 1])) ? 1 : 0;
    for (int o = kWave / 2; o > 0; o >>= 1) c += __shfl_down(c, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd(count, c);
}

struct DevBuf {
    std::vector<void*> v;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { clear(); }
    void clear() {
        for (void* p : v) cache_free(p);
        v.clear();
    }
    template <typename T>
    int get(T** p, size_t count) {
        int e = talloc(p, count);
        if (!e) v.push_back((void*)*p);
        return e;
    }
    void release(void* p) {
        auto it = std::find(v.begin(), v.end(), p);
        if (it != v.end()) {
            cache_free(p);
            v.erase(it);
        }
    }
};

template <typename TX>
int gram_cols(const TX* const* cols, int C, int64_t n, DevBuf& tmp, std::vector<double>& G) {
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(kMeatBlocks, (n + kBlock - 1) / kBlock));
    double* part = nullptr;
    int e = tmp.get(&part, (size_t)grid * (16 * 17 / 2));
    if (e) return e;
    hipLaunchKernelGGL((gram_cols_kernel<TX>), dim3(grid), dim3(kBlock), 0, 0, cols, C, n, part);
    FEHIP_LAUNCH_CHECK("gram (matrix cores)");
    return reduce_upper(part, (int)grid, 16, C, G);
}

// Device array of column pointers (the kernels that read [y W Z] take the columns through it, so a
// request's columns may live in its own work buffer or in storage shared with other requests).
template <typename TX>
int column_array(DevBuf& db, const std::vector<const TX*>& cols, const TX* const** out) {
    const TX** d = nullptr;
    int e = db.get(&d, std::max<size_t>(1, cols.size()));
    if (e) return e;
    FEHIP_CHECK(copy_sync(d, cols.data(), sizeof(const TX*) * cols.size(), hipMemcpyHostToDevice), "column pointers");
    *out = d;
    return 0;
}

int host_sum_parts(const double* dpart, int nblk, int width, std::vector<double>& out) {
    std::vector<double> h((size_t)nblk * width);
    FEHIP_CHECK(copy_sync(h.data(), dpart, h.size() * sizeof(double), hipMemcpyDeviceToHost), "partials");
    out.assign(width, 0.0);
    for (int b = 0; b < nblk; ++b)
        for (int k = 0; k < width; ++k) out[k] += h[(size_t)b * width + k];
    return 0;
}

int sumsq(const double* a, int64_t n, double* part, double* out) {
    hipLaunchKernelGGL(dot_part_kernel, dim3(kRedBlocks), dim3(kBlock), 0, 0, a, n, part);
    FEHIP_LAUNCH_CHECK("sumsq");
    std::vector<double> s;
    int e = host_sum_parts(part, kRedBlocks, 1, s);
    if (e) return e;
    *out = s[0];
    return 0;
}

void finalize_stats(fehip_result* r, int64_t n_obs) {
    const double EPS = 1e-12;
    const double n = (double)std::max<int64_t>(1, n_obs);
    const double dfm = (double)std::max(0, r->dof_model);
    const double dfr = (double)std::max<int64_t>(1, r->dof_resid);
    const double dev = r->rss;
    const double dev0 = std::max(EPS, r->tss_total);
    r->r2 = dev0 > EPS ? std::max(0.0, std::min(1.0, 1.0 - dev / dev0)) : 0.0;
    r->r2_within = r->tss_within > EPS ? std::max(0.0, std::min(1.0, 1.0 - dev / r->tss_within)) : r->r2;
    const double k_total = (double)(r->dof_model + r->dof_fes + 1);
    const double denom = std::max(n - k_total, 1.0);
    r->r2_adj = (n > 1.0 && dev0 > EPS) ? std::max(0.0, std::min(1.0, 1.0 - (dev * (n - 1.0)) / (dev0 * denom))) : r->r2;
    const double ssr = std::max(0.0, dev0 - dev);
    r->f_stat = (dfm > 0.0 && dfr > 0.0 && dev > EPS) ? (ssr / dfm) / (dev / dfr) : 0.0;
}


// How clusters of a subset containing the primary lie inside the segments (cached per panel): compares
// the subset's cluster count with its count of runs inside segments.
int nest_mode(Panel* p, unsigned dims_mask, const int* const* kid, int nk, unsigned long long* dcount, int* mode) {
    std::lock_guard<std::recursive_mutex> guard(p->mu);
    auto it = p->nest_modes.find(dims_mask);
    if (it != p->nest_modes.end()) {
        *mode = it->second;
        return 0;
    }
    int m = 1;
    if (nk > 0) {
        const Partition* part_s = nullptr;
        int e = panel_partition(p, dims_mask, &part_s, 0);
        if (e) return e;
        FEHIP_CHECK(hipMemsetAsync(dcount, 0, sizeof(unsigned long long), 0), "memset");
        hipLaunchKernelGGL(composite_runs_kernel, dim3(stream_grid((size_t)p->n, kBlock)), dim3(kBlock), 0, 0,
                           (const int*)p->d[p->P].ids, kid[0], nk > 1 ? kid[1] : nullptr, p->n, dcount);
        FEHIP_LAUNCH_CHECK("composite runs");
        unsigned long long runs = 0;
        FEHIP_CHECK(copy_sync(&runs, dcount, sizeof(runs), hipMemcpyDeviceToHost), "runs copy");
        m = part_s->n_clusters == p->n ? 2 : ((unsigned long long)part_s->n_clusters == runs ? 1 : 0);
    }
    p->nest_modes[dims_mask] = m;
    *mode = m;
    return 0;
}

// Mx = proj' Mz proj for the 2SLS meat (Mz: m x m, proj: m x k, column-major); small on the host,
// wide on the device.
int project_meat(const std::vector<double>& Mz, int m, const std::vector<double>& proj, int k, std::vector<double>& Mx) {
    Mx.assign((size_t)k * k, 0.0);
    if ((int64_t)m * m * k <= (int64_t)4 << 20) {
        std::vector<double> t((size_t)m * k, 0.0);  // Mz proj
        for (int j = 0; j < k; ++j)
            for (int b = 0; b < m; ++b) {
                const double pb = proj[(size_t)j * m + b];
                if (pb == 0.0) continue;
                for (int a = 0; a < m; ++a) t[(size_t)j * m + a] += Mz[(size_t)b * m + a] * pb;
            }
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < k; ++i) {
                double v = 0.0;
                for (int a = 0; a < m; ++a) v += proj[(size_t)i * m + a] * t[(size_t)j * m + a];
                Mx[(size_t)j * k + i] = v;
            }
        return 0;
    }
    double *dM = nullptr, *dP = nullptr, *dT = nullptr, *dX = nullptr;
    int e;
    if ((e = talloc(&dM, (size_t)m * m)) || (e = talloc(&dP, (size_t)m * k)) || (e = talloc(&dT, (size_t)m * k)) ||
        (e = talloc(&dX, (size_t)k * k)))
        return e;
    struct Free {
        double** v[4];
        ~Free() {
            for (auto p : v) cache_free(*p);
        }
    } fr{{&dM, &dP, &dT, &dX}};
    FEHIP_CHECK(copy_sync(dM, Mz.data(), sizeof(double) * (size_t)m * m, hipMemcpyHostToDevice), "meat copy");
    FEHIP_CHECK(copy_sync(dP, proj.data(), sizeof(double) * (size_t)m * k, hipMemcpyHostToDevice), "proj copy");
    if ((e = gemm_impl(false, false, m, k, m, 1.0, dM, m, dP, m, 0.0, dT, m, false))) return e;
    if ((e = gemm_impl(true, false, k, k, m, 1.0, dP, m, dT, m, 0.0, dX, k, false))) return e;
    FEHIP_CHECK(copy_sync(Mx.data(), dX, sizeof(double) * (size_t)k * k, hipMemcpyDeviceToHost), "meat copy");
    return 0;
}

// Which subsets of the sandwich the fused pass over the wave tiles forms, and how: the slot layout of
// the other cluster dims and the pass's output tables.  The rest go to the methods that read the
// residual or the score matrix.
enum {
    kRouteOther = 0,
    kRouteNest = 1,
    kRouteRun = 2,
    kRouteSmall = 3,
    kRouteAlias = 4,
    kRouteSorted = 5,
    kRouteRunCells = 6,
    kRouteWide = 7
};

struct ScorePlan {
    bool fused = false;   // a fused pass runs
    int KT = 0;
    WaveScores ws{};      // slots and output tables (buffers are set at the launch)
    int route[8] = {};    // per subset mask
    int index[8] = {};    // nested subset, run output or small table of the subset (alias: the subset it repeats)
    int run_dim[kMaxRunOuts] = {};
    int small_cells[kMaxSmallOuts] = {};
    bool needs_r = false;  // a subset left to the other methods reads the residual or the score matrix
    bool reads_T = false;  // ... the score matrix
    int sorted_mask = 0;   // the subset whose score rows the pass writes in a partition's order (0: none)
    unsigned sorted_dims = 0;  // the dims of that partition (the subset's, or its large dim's for RunCells)
    int cells = 0;             // RunCells: cells of the subset's small dims
    // Wide: subsets of the primary and one large dim D read the segment sums and D's run sums of the
    // wide pass
    int wide_dim = -1;
    bool wide_seg = false, wide_run = false;
};

int plan_scores(Panel* p, const int* cdims, int ncd, int Kt, unsigned long long* dcount, ScorePlan& pl) {
    pl = ScorePlan();
    const bool fused_ok = Kt <= 12 && p->n_long == 0 && p->n_mtiles == 0 && p->n_wtiles > 0;
    pl.KT = Kt <= 4 ? 4 : (Kt <= 8 ? 8 : 12);
    WaveScores& w = pl.ws;
    auto slot_of = [&](int d) {
        for (int j = 0; j < 3; ++j)
            if (w.kd[j] == p->d[d].ids) return j;
        for (int j = 0; j < 3; ++j)
            if (!w.kd[j]) {
                w.kd[j] = p->d[d].ids;
                return j;
            }
        return -1;  // not reached: at most three cluster dims
    };
    int single_q = -1;
    // scores too wide for the fused pass: the subsets of the primary and the first large cluster dim
    // go through the wide pass, which needs the tiles of whole segments to cover the panel
    unsigned wide_ok = 0;
    int wide_d = -1;
    if (Kt > 12 && p->n_long == 0 && p->n_tiles > 0) {
        for (int j = 0; j < ncd && wide_d < 0; ++j) {
            const int d = cdims[j] - 1;
            if (d != p->P && !p->d[d].small && p->d[d].n_runs > 0) wide_d = d;
        }
        wide_ok = (1u << p->P) | (wide_d >= 0 ? 1u << wide_d : 0u);
    }
    for (int mask = 1; mask < (1 << ncd); ++mask) {
        unsigned dm = 0;
        int len = 0;
        for (int j = 0; j < ncd; ++j)
            if (mask & (1 << j)) {
                dm |= 1u << (cdims[j] - 1);
                ++len;
            }
        const bool has_p = (dm >> p->P) & 1u;
        bool all_small = true;
        int64_t cells = 1;
        for (int d = 0; d < p->m; ++d)
            if ((dm >> d) & 1u) {
                all_small &= p->d[d].small != 0;
                cells *= p->d[d].G;
            }
        const int single = len == 1 ? __builtin_ctz(dm) : -1;
        int route = kRouteOther;
        if (fused_ok && has_p) {
            const int* kid[2] = {nullptr, nullptr};
            int nk = 0;
            for (int d = 0; d < p->m; ++d)
                if (d != p->P && ((dm >> d) & 1u)) kid[nk++] = p->d[d].ids;
            int mode = 1;
            int e = nest_mode(p, dm, kid, nk, dcount, &mode);
            if (e) return e;
            if (mode == 2 && single_q >= 0) {  // singletons: the meat of the first such subset
                route = kRouteAlias;
                pl.index[mask] = single_q;
            } else if (w.nsub < kMaxNestSubsets) {
                int use = 0;
                for (int d = 0; d < p->m; ++d)
                    if (d != p->P && ((dm >> d) & 1u)) use |= 1 << slot_of(d);
                if (mode == 2) single_q = mask;
                w.use[w.nsub] = use;
                w.mode[w.nsub] = mode;
                pl.index[mask] = w.nsub++;
                route = kRouteNest;
            }
        } else if (fused_ok && single >= 0 && single != p->P && !p->d[single].small && w.nrun < kMaxRunOuts) {
            const int j = w.nrun++;
            w.rslot[j] = slot_of(single);
            w.rshare[j] = -1;
            w.rdest[j] = p->d[single].row_rdest;
            pl.run_dim[j] = single;
            pl.index[mask] = j;
            route = kRouteRun;
        } else if (fused_ok && all_small && w.nsmall < kMaxSmallOuts &&
                   (int64_t)w.small_lds + cells * (pl.KT + 1) <= kScoreSmallLds) {
            const int q = w.nsmall++;
            int stride = 1;
            for (int j = 0; j < 3; ++j) w.sstride[q][j] = 0;
            for (int d = 0; d < p->m; ++d)
                if ((dm >> d) & 1u) {
                    w.sstride[q][slot_of(d)] = stride;
                    stride *= p->d[d].G;
                }
            w.soff[q] = w.small_lds;
            w.small_lds += (int)cells * (pl.KT + 1);
            pl.small_cells[q] = (int)cells;
            pl.index[mask] = q;
            route = kRouteSmall;
        }
        if (route == kRouteOther && fused_ok && !has_p && pl.sorted_mask == 0) {
            // one subset through a sorted partition: the pass writes its score rows in that order, so the
            // partition kernel reads them in sequence instead of gathering them.  A large dim crossed with
            // small dims of few cells takes the large dim's partition, whose runs stay contiguous, and the
            // rows' cells; any other such subset takes its own partition.
            int large = -1, nlarge = 0;
            int64_t bcells = 1;
            for (int d = 0; d < p->m; ++d)
                if ((dm >> d) & 1u) {
                    if (p->d[d].small) {
                        bcells *= p->d[d].G;
                    } else {
                        large = d;
                        ++nlarge;
                    }
                }
            pl.sorted_mask = mask;
            if (nlarge == 1 && len > 1 && bcells <= kMaxRunCells) {
                pl.sorted_dims = 1u << large;
                pl.cells = (int)bcells;
                int stride = 1;
                for (int d = 0; d < p->m; ++d)
                    if (((dm >> d) & 1u) && d != large) {
                        w.cstride[slot_of(d)] = stride;
                        stride *= p->d[d].G;
                    }
                route = kRouteRunCells;
            } else {
                pl.sorted_dims = dm;
                route = kRouteSorted;
            }
        }
        if (route == kRouteOther && wide_ok && (dm & ~wide_ok) == 0) {
            route = kRouteWide;
            pl.wide_dim = wide_d;
            pl.wide_seg |= has_p;
            pl.wide_run |= wide_d >= 0 && ((dm >> wide_d) & 1u);
            pl.needs_r = true;
        }
        if (route == kRouteOther) {
            // the other methods: block-tile nested, small tables and run sums read the residual; the
            // block-tile nested, sorted-partition and segment-sum methods read the score matrix
            bool reads_T = true;
            if (has_p && p->n_long == 0 && Kt <= 12)
                reads_T = p->n_mtiles != 0;
            else if (all_small && cells <= kSmallCells && Kt <= 11)
                reads_T = false;
            else if (single >= 0 && single != p->P && !p->d[single].small && p->n_long == 0 && Kt <= 12)
                reads_T = false;
            pl.reads_T |= reads_T;
            pl.needs_r = true;
        }
        pl.route[mask] = route;
    }
    // a run dim whose runs are those of a nested subset of the primary and that dim alone shares its scan
    for (int j = 0; j < w.nrun; ++j)
        for (int q = 0; q < w.nsub; ++q)
            if (w.use[q] == (1 << w.rslot[j])) w.rshare[j] = q;
    pl.fused = w.nsub > 0 || w.nrun > 0 || w.nsmall > 0 || pl.sorted_mask != 0;
    return 0;
}

// One request of a batch: its inputs, and what the preparation step leaves for the solve and the
// normal equations / sandwich step.
struct ReqState {
    const double* d_y = nullptr;
    const double* d_W = nullptr;
    const double* d_Z = nullptr;
    int K = 0, L = 0;
    int64_t ld = 0;
    const uint8_t* d_weights = nullptr;
    const fehip_opts* opts = nullptr;
    fehip_result* result = nullptr;
    double *beta = nullptr, *se = nullptr, *vcov = nullptr;
    // prepared
    unsigned fe_mask = 0;
    int ncd = 0;
    int cdims[3] = {0, 0, 0};
    DevBuf db;
    double* part = nullptr;
    uint8_t* wpanel = nullptr;
    int64_t N = 0;
    int Geff[FEHIP_MAX_FE] = {0};
    int dof_fes = 0, dof_nested = 0;
    int C = 0;
    double* work = nullptr;    // n x C within-transformed [y W Z], panel order (FP64 mode)
    float* work32 = nullptr;   // the same stored in FP32 (mixed mode)
    double* alpha = nullptr;   // mixed mode: solved FE tables of every column, [C][S]
    int64_t S = 0;
    // FP64 cross products of the C transformed columns (C x C, column-major) formed in the final pass
    // of the within-transform, when all of the request's columns were finalized in one column tile
    std::vector<double> gram;
    // columns held in storage shared by the requests of a batch (no work buffer of its own): the
    // device pointer of each of the C columns (FP32 when work32 would be used, else FP64)
    std::vector<const void*> shared_cols;
    bool shared = false;  // set before preparation: no work buffer of its own
    bool f32 = false;     // transformed columns stored in FP32
    clk::time_point t0, t1;
};

// Per-batch results that depend only on the outcome column and the row subsample: N and the total
// sum of squares of y (two passes over y), and the non-empty group counts of a subsample.
struct BatchShared {
    std::map<std::pair<const double*, const uint8_t*>, std::pair<int64_t, double>> ysums;
    std::map<const uint8_t*, std::vector<int>> geff;
};

// Validation, cluster dims, weights in panel order, N, TSS, effective group counts and DoF, and
// the work buffer of the request.
int prepare_request(Panel* p, ReqState& s, BatchShared& bs) {
    const int64_t n = p->n;
    const int K = s.K, L = s.L;
    const int64_t ld = s.ld;
    const double *d_y = s.d_y, *d_W = s.d_W, *d_Z = s.d_Z;
    const uint8_t* d_weights = s.d_weights;
    const fehip_opts* opts = s.opts;
    fehip_result* result = s.result;
    double *beta = s.beta, *se = s.se, *vcov = s.vcov;
    (void)d_W;
    (void)d_Z;
    if (!opts || !result || !d_y || (K > 0 && (!beta || !se))) return set_error("fehip_estimate: null argument");
    if (K < 0 || L < 0 || (K > 0 && !d_W) || (L > 0 && !d_Z) || ld < n) return set_error("fehip_estimate: bad columns");
    std::memset(result, 0, sizeof(*result));
    for (int k = 0; k < K; ++k) {
        beta[k] = 0.0;
        se[k] = 0.0;
    }
    if (vcov) std::fill(vcov, vcov + (size_t)K * K, 0.0);
    DevBuf& db = s.db;
    int e;
    const auto t0 = clk::now();
    const unsigned fe_mask = (opts->demean.fe_mask ? opts->demean.fe_mask : ((1u << p->m) - 1u)) & ((1u << p->m) - 1u);
    // cluster dims: valid, unique, first 3 in request order (initialize_cluster_dims)
    int ncd = 0;
    int cdims[3];
    for (int i = 0; i < opts->n_cluster && i < 4; ++i) {
        const int d = opts->cluster_dims[i];
        if (d < 1 || d > p->m) continue;
        bool dup = false;
        for (int j = 0; j < ncd; ++j) dup |= cdims[j] == d;
        if (dup) continue;
        if (ncd == 3) break;
        cdims[ncd++] = d;
    }
    result->n_cluster_dims = ncd;
    for (int j = 0; j < ncd; ++j) result->cluster_dims[j] = cdims[j];

    double* part = nullptr;
    if ((e = db.get(&part, (size_t)std::max(kRedBlocks * 3, kResidBlocks)))) return e;
    // weights in panel order, effective N, TSS of y (compute_tss with intercept)
    uint8_t* wpanel = nullptr;
    if (d_weights) {
        if ((e = db.get(&wpanel, (size_t)n))) return e;
        hipLaunchKernelGGL(weights_panel_kernel, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, 0, d_weights, p->perm, n,
                           wpanel);
    }
    const auto ykey = std::make_pair(d_y, d_weights);
    auto yit = bs.ysums.find(ykey);
    if (yit == bs.ysums.end()) {
        std::vector<double> s3;
        hipLaunchKernelGGL(wsum2_kernel, dim3(kRedBlocks), dim3(kBlock), 0, 0, d_y, d_weights, n, 0.0, part);
        if ((e = host_sum_parts(part, kRedBlocks, 3, s3))) return e;
        const int64_t N0 = (int64_t)llround(s3[2]);
        const double ybar = N0 > 0 ? s3[0] / (double)N0 : 0.0;
        hipLaunchKernelGGL(wsum2_kernel, dim3(kRedBlocks), dim3(kBlock), 0, 0, d_y, d_weights, n, ybar, part);
        if ((e = host_sum_parts(part, kRedBlocks, 3, s3))) return e;
        yit = bs.ysums.emplace(ykey, std::make_pair(N0, s3[1])).first;
    }
    const int64_t N = yit->second.first;
    result->n_obs = N;
    result->tss_total = yit->second.second;
    // group counts: G_d (max id) without weights, non-empty groups with weights
    int Geff[FEHIP_MAX_FE];
    auto git = wpanel ? bs.geff.find(d_weights) : bs.geff.end();  // unweighted: G_d = max id
    if (wpanel && git == bs.geff.end()) {  // every dim in one pass and one read-back
        DimIds dk{};
        dk.m = p->m;
        dk.off[0] = 0;
        for (int d = 0; d < p->m; ++d) {
            dk.ids[d] = p->d[d].ids;
            dk.off[d + 1] = dk.off[d] + p->d[d].G;
        }
        int* flags = nullptr;
        unsigned long long* cnt = nullptr;
        if ((e = db.get(&flags, (size_t)dk.off[p->m])) || (e = db.get(&cnt, FEHIP_MAX_FE))) return e;
        FEHIP_CHECK(hipMemsetAsync(flags, 0, sizeof(int) * (size_t)dk.off[p->m], 0), "memset");
        FEHIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * FEHIP_MAX_FE, 0), "memset");
        hipLaunchKernelGGL(nonempty_dims_kernel, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, 0, dk,
                           (const uint8_t*)wpanel, n, flags);
        hipLaunchKernelGGL(count_dims_kernel, dim3(stream_grid((size_t)dk.off[p->m], kBlock)), dim3(kBlock), 0, 0,
                           (const int*)flags, dk, cnt);
        FEHIP_LAUNCH_CHECK("group counts");
        unsigned long long h[FEHIP_MAX_FE] = {0};
        FEHIP_CHECK(copy_sync(h, cnt, sizeof(h), hipMemcpyDeviceToHost), "count");
        bs.geff[d_weights] = std::vector<int>(h, h + p->m);
        git = bs.geff.find(d_weights);
    }
    for (int d = 0; d < p->m; ++d) Geff[d] = git != bs.geff.end() ? git->second[d] : p->d[d].G;
    int dof_fes = 0, dof_nested = 0;
    for (int d = 0; d < p->m; ++d)
        if ((fe_mask >> d) & 1u) dof_fes += std::max(0, Geff[d] - 1);
    for (int j = 0; j < ncd; ++j) dof_nested += std::max(0, Geff[cdims[j] - 1] - 1);
    result->dof_fes = dof_fes;
    s.fe_mask = fe_mask;
    s.ncd = ncd;
    for (int j = 0; j < ncd; ++j) s.cdims[j] = cdims[j];
    s.part = part;
    s.wpanel = wpanel;
    s.N = N;
    for (int d = 0; d < p->m; ++d) s.Geff[d] = Geff[d];
    s.dof_fes = dof_fes;
    s.dof_nested = dof_nested;
    s.t0 = t0;
    s.C = 1 + K + L;
    // 2SLS keeps FP64 storage in the mixed mode: the first-stage projection (Z'Z, Z'X) must be FP64,
    // and an over-identified fit does not drive Z'e to zero, so FP32 cross products would bias beta
    const bool iv_request = opts->n_iv_reg > 0 && L > 0;
    if (opts->demean.precision != 0 && opts->demean.precision != 1)
        return set_error("fehip_estimate: precision must be 0 (FP64) or 1 (FP32 storage)");
    s.f32 = opts->demean.precision == 1 && !iv_request;
    if (s.shared) return 0;  // the batch holds the columns
    if (s.f32) {
        if ((e = db.get(&s.work32, (size_t)n * s.C))) return e;
        s.S = within_table_size(p, fe_mask);
        if (s.S < 0) return 1;
        if (s.S > 0 && (e = db.get(&s.alpha, (size_t)s.S * s.C))) return e;
    } else {
        if ((e = db.get(&s.work, (size_t)n * s.C))) return e;
    }
    return 0;
}

// Normal equations, IV stages, column filter, Cholesky, residual and the sandwich of a request whose
// work buffer holds the within-transformed columns.
int finish_request(Panel* p, ReqState& s) {
    const int64_t n = p->n;
    const int K = s.K, L = s.L;
    const int64_t ld = s.ld;
    const double *d_y = s.d_y, *d_W = s.d_W, *d_Z = s.d_Z;
    const uint8_t* d_weights = s.d_weights;
    const fehip_opts* opts = s.opts;
    fehip_result* result = s.result;
    double *beta = s.beta, *se = s.se, *vcov = s.vcov;
    (void)d_Z;
    DevBuf& db = s.db;
    int e;
    const unsigned fe_mask = s.fe_mask;
    (void)fe_mask;
    const int ncd = s.ncd;
    const int* cdims = s.cdims;
    double* part = s.part;
    uint8_t* wpanel = s.wpanel;
    const int64_t N = s.N;
    const int* Geff = s.Geff;
    const int dof_fes = s.dof_fes, dof_nested = s.dof_nested;
    const int C = s.C;
    double* work = s.work;
    const auto t1 = s.t1;
    if (K <= 0) return 0;
    // the request's columns [y~ W~ Z~] through a device pointer array
    const double* const* dcols = nullptr;
    const float* const* dcols32 = nullptr;
    const void* ycol = nullptr;  // y~
    {
        if (s.f32) {
            std::vector<const float*> hp(C);
            for (int j = 0; j < C; ++j)
                hp[j] = s.shared_cols.empty() ? s.work32 + (int64_t)j * n : (const float*)s.shared_cols[j];
            if ((e = column_array(db, hp, &dcols32))) return e;
            ycol = hp[0];
        } else {
            std::vector<const double*> hp(C);
            for (int j = 0; j < C; ++j)
                hp[j] = s.shared_cols.empty() ? s.work + (int64_t)j * n : (const double*)s.shared_cols[j];
            if ((e = column_array(db, hp, &dcols))) return e;
            ycol = hp[0];
        }
    }

    // ---- Gram of [y W Z] (upper), host copy
    double* mpart = nullptr;
    if ((e = db.get(&mpart, (size_t)kMeatBlocks * 78))) return e;
    std::vector<double> G((size_t)C * C);
    const bool f32 = s.f32;
    float* work32 = s.work32;
    const bool have_gram = s.gram.size() == (size_t)C * C;
    if (have_gram) {
        G = s.gram;
    } else if (C > 12 && C <= 16) {
        if (f32 ? (e = gram_cols(dcols32, C, n, db, G)) : (e = gram_cols(dcols, C, n, db, G))) return e;
    } else if (C <= 12) {
        if (f32) {
            if ((e = gram_fused(work32, n, C, n, mpart, G))) return e;
        } else if ((e = gram_fused((const double*)work, n, C, n, mpart, G))) {
            return e;
        }
    } else {
        double* Gd = nullptr;
        if ((e = db.get(&Gd, (size_t)C * C))) return e;
        const double* src = work;
        if (f32) {  // wide designs: the split-K GEMM reads FP64
            double* wide = nullptr;
            if ((e = db.get(&wide, (size_t)n * C))) return e;
            hipLaunchKernelGGL(widen_kernel, dim3(stream_grid((size_t)n * C, kBlock)), dim3(kBlock), 0, 0,
                               (const float*)work32, n * C, wide);
            src = wide;
        }
        if ((e = gemm_impl(true, false, C, C, n, 1.0, src, n, src, n, 0.0, Gd, C, true))) return e;
        FEHIP_CHECK(copy_sync(G.data(), Gd, G.size() * sizeof(double), hipMemcpyDeviceToHost), "gram copy");
        for (int j = 0; j < C; ++j)
            for (int i = j + 1; i < C; ++i) G[(size_t)j * C + i] = G[(size_t)i * C + j];
    }
    auto g = [&](int i, int j) { return G[(size_t)j * C + i]; };
    result->tss_within = g(0, 0);

    // endogenous / exogenous / instruments (fe_pipeline.f90:664-753)
    std::vector<int> is_endog(K, 0);
    for (int i = 0; i < opts->n_iv_reg; ++i) {
        const int v = opts->iv_reg[i];
        if (v >= 1 && v <= K) is_endog[v - 1] = 1;
    }
    std::vector<int> exog, inst;
    int n_endog = 0;
    for (int k = 0; k < K; ++k) {
        if (is_endog[k])
            ++n_endog;
        else
            exog.push_back(k);
    }
    if (L > 0) {
        for (int i = 0; i < opts->n_iv_inst; ++i)
            if (opts->iv_inst[i] >= 1 && opts->iv_inst[i] <= L) inst.push_back(opts->iv_inst[i] - 1);
        if (inst.empty())
            for (int l = 0; l < L; ++l) inst.push_back(l);
    }
    bool use_iv = n_endog > 0 && (int)inst.size() >= n_endog;
    const int n_aug = (int)exog.size() + (int)inst.size();
    result->is_iv = use_iv;
    result->n_instruments = use_iv ? n_aug : 0;
    std::vector<double> Q((size_t)K * K), b(K);
    std::vector<double> proj;  // n_aug x K
    std::vector<int> aug_col;  // work column of each Z_aug column
    if (use_iv) {
        for (int k : exog) aug_col.push_back(1 + k);
        for (int l : inst) aug_col.push_back(1 + K + l);
        std::vector<double> Qzz((size_t)n_aug * n_aug), Qzy(n_aug), Qzx((size_t)n_aug * K);
        for (int a = 0; a < n_aug; ++a) {
            Qzy[a] = g(aug_col[a], 0);
            for (int c = 0; c < n_aug; ++c) Qzz[(size_t)c * n_aug + a] = g(aug_col[a], aug_col[c]);
            for (int k = 0; k < K; ++k) Qzx[(size_t)k * n_aug + a] = g(aug_col[a], 1 + k);
        }
        std::vector<double> rhs, Qzinv;
        int info = dense::chol_solve_invert(Qzz, Qzy, n_aug, rhs, Qzinv);
        if (info != 0) {
            result->solver_info = info;
            return 0;
        }
        proj.assign((size_t)n_aug * K, 0.0);
        for (int k = 0; k < K; ++k)
            for (int a = 0; a < n_aug; ++a) {
                double s = 0.0;
                for (int c = 0; c < n_aug; ++c) s += Qzinv[(size_t)c * n_aug + a] * Qzx[(size_t)k * n_aug + c];
                proj[(size_t)k * n_aug + a] = s;
            }
        for (int i = 0; i < K; ++i)
            for (int j = 0; j < K; ++j) {
                double s = 0.0;
                for (int a = 0; a < n_aug; ++a) s += Qzx[(size_t)i * n_aug + a] * proj[(size_t)j * n_aug + a];
                Q[(size_t)j * K + i] = s;
            }
        for (int i = 0; i < K; ++i)
            for (int j = i + 1; j < K; ++j) {
                const double s = 0.5 * (Q[(size_t)j * K + i] + Q[(size_t)i * K + j]);
                Q[(size_t)j * K + i] = Q[(size_t)i * K + j] = s;
            }
        for (int i = 0; i < K; ++i) {
            double s = 0.0;
            for (int a = 0; a < n_aug; ++a) s += Qzx[(size_t)i * n_aug + a] * rhs[a];
            b[i] = s;
        }
    } else {
        for (int i = 0; i < K; ++i) {
            b[i] = g(1 + i, 0);
            for (int j = 0; j < K; ++j) Q[(size_t)j * K + i] = g(1 + i, 1 + j);
        }
    }
    // column filter on diag(W~'W~) and Cholesky (fe_pipeline.f90:1034-1107)
    double dmax = 0.0;
    for (int k = 0; k < K; ++k) dmax = std::max(dmax, std::fabs(g(1 + k, 1 + k)));
    dmax = std::max(dmax, 1.0);
    const double dtol = 1e-10 * dmax;
    std::vector<int> keep;
    for (int k = 0; k < K; ++k)
        if (!(std::fabs(g(1 + k, 1 + k)) <= dtol)) keep.push_back(k);
    const int kept = (int)keep.size();
    result->dof_model = kept;
    if (kept == 0) {
        result->solver_info = -1;
        return 0;
    }
    std::vector<double> Qs((size_t)kept * kept), bs(kept), bsol, Qinv;
    for (int a = 0; a < kept; ++a) {
        bs[a] = b[keep[a]];
        for (int c = 0; c < kept; ++c) Qs[(size_t)c * kept + a] = Q[(size_t)keep[c] * K + keep[a]];
    }
    int info = dense::chol_solve_invert(Qs, bs, kept, bsol, Qinv);
    result->solver_info = info;
    if (info != 0) return 0;
    for (int a = 0; a < kept; ++a) beta[keep[a]] = bsol[a];
    result->rank = kept;
    const auto t2 = clk::now();
    result->t_regression = std::chrono::duration<double>(t2 - t1).count();

    // ---- residual on W~ (also for IV), rss, dof
    double *dbeta = nullptr, *r = nullptr;
    if ((e = db.get(&dbeta, (size_t)K)) || (e = db.get(&r, (size_t)n))) return e;
    ScorePlan plan;
    if (ncd > 0) {
        unsigned long long* pc = nullptr;
        if ((e = db.get(&pc, 1)) || (e = plan_scores(p, cdims, ncd, use_iv ? n_aug : K, pc, plan))) return e;
    }
    // the fused pass forms the residual itself when no other method reads it (OLS only: the 2SLS scores
    // are formed from other columns than the residual's)
    bool resid_in_pass = false;
    // score rows T = r W~ are formed in the residual pass when an OLS sandwich has a subset whose method
    // reads them (every subset but those containing the primary on wave tiles)
    double* T_early = nullptr;
    const double* const* iv_scores = nullptr;  // 2SLS: the instrument-set columns the scores are formed from
    bool rss_done = false;
    double rss = 0.0;
    if (have_gram && ncd == 0 && !use_iv) {
        // no sandwich: rss from the FP64 cross products, no pass over the rows
        double bq = 0.0, by = 0.0;
        for (int i = 0; i < K; ++i) {
            by += beta[i] * g(1 + i, 0);
            for (int j = 0; j < K; ++j) bq += beta[i] * g(1 + i, 1 + j) * beta[j];
        }
        rss = std::max(0.0, g(0, 0) - 2.0 * by + bq);
        rss_done = true;
        // y~'y~ - 2 b'W~'y~ + b'W~'W~b loses about eps * y~'y~ to cancellation: when the fit leaves less
        // than 1e-6 of y~'y~ (within R^2 near 1), FP64 columns take rss from the residual itself, as the
        // reference does (r = y~ - W~b, then r'r).  FP32 columns keep the FP64 cross products, whose
        // error is below that of a residual formed from the rounded columns.
        if (!f32 && !(rss > 1e-6 * g(0, 0))) {
            FEHIP_CHECK(copy_sync(dbeta, beta, sizeof(double) * K, hipMemcpyHostToDevice), "beta copy");
            hipLaunchKernelGGL(resid_scores_kernel<double>, dim3(kResidBlocks), dim3(256), 0, 0, dcols, n, K,
                               (const double*)dbeta, r, (double*)nullptr, K, dcols + 1, part);
            FEHIP_LAUNCH_CHECK("residual");
            std::vector<double> ps;
            if ((e = host_sum_parts(part, kResidBlocks, 1, ps))) return e;
            rss = ps[0];
        }
    } else if (!f32 || have_gram) {
        // With FP32 storage and cross products formed in FP64 before the rounding, beta above is already
        // the FP64 solution (no refinement); the residual and scores then read the FP32 columns, whose
        // rounding enters the sandwich only at the level of the storage precision, and rss comes from
        // the FP64 cross products.
        // Scores: r W~, or for 2SLS r Z_aug with Z_aug the exogenous regressors and the instruments, read in
        // place from the work columns.  The sandwich's subset methods that read the score matrix get it
        // from this pass; when the fused pass over the wave tiles forms every subset of an OLS sandwich,
        // that pass forms the residual itself and this one is skipped.
        const int Kt = use_iv ? n_aug : K;
        if (use_iv && ncd > 0) {
            std::vector<const double*> zp(n_aug);
            for (int a2 = 0; a2 < n_aug; ++a2) zp[a2] = work + (int64_t)aug_col[a2] * n;
            if ((e = column_array(db, zp, &iv_scores))) return e;
        }
        FEHIP_CHECK(copy_sync(dbeta, beta, sizeof(double) * K, hipMemcpyHostToDevice), "beta copy");
        if (ncd > 0 && !use_iv && plan.fused && !plan.needs_r) {
            resid_in_pass = true;  // rss: below from the cross products (FP32 storage) or from the pass
        } else {
            if (ncd > 0 && plan.reads_T && (e = db.get(&T_early, (size_t)n * Kt))) return e;
            const size_t tlds = (T_early && Kt <= 16) ? sizeof(double) * 256 * (size_t)Kt : 0;
            if (f32)
                hipLaunchKernelGGL(resid_scores_kernel<float>, dim3(kResidBlocks), dim3(256), tlds, 0, dcols32, n, K,
                                   (const double*)dbeta, r, T_early, Kt, dcols32 + 1, part);
            else
                hipLaunchKernelGGL(resid_scores_kernel<double>, dim3(kResidBlocks), dim3(256), tlds, 0, dcols, n, K,
                                   (const double*)dbeta, r, T_early, Kt, use_iv ? iv_scores : dcols + 1, part);
            FEHIP_LAUNCH_CHECK("residual");
            std::vector<double> ps;
            if ((e = host_sum_parts(part, kResidBlocks, 1, ps))) return e;
            rss = ps[0];
        }
        if (f32) {  // y~'y~ - 2 beta' W~'y~ + beta' W~'W~ beta from the FP64 cross products
            double bq = 0.0, by = 0.0;
            for (int i = 0; i < K; ++i) {
                by += beta[i] * g(1 + i, 0);
                for (int j = 0; j < K; ++j) bq += beta[i] * g(1 + i, 1 + j) * beta[j];
            }
            rss = std::max(0.0, g(0, 0) - 2.0 * by + bq);
        }
        rss_done = true;
    } else {
        // FP64 refinement on the residual: e = M_F(y - W beta) in FP64 from the caller's FP64 columns and
        // the same combination of the solved tables (M_F is linear, so no further solve), then
        // beta += Q^-1 X~'e; one sweep (the correction's own error is second order in the FP32
        // rounding).  The refined residual feeds rss and the sandwich.
        double *ec = nullptr, *et = nullptr, *ae = nullptr, *xpart = nullptr;
        uint32_t* wm1 = nullptr;
        const int kRefineBlocks = 256;
        if ((e = db.get(&ec, (size_t)n)) || (e = db.get(&et, (size_t)n)) ||
            (e = db.get(&ae, (size_t)std::max<int64_t>(1, s.S))) || (e = db.get(&xpart, (size_t)kRefineBlocks * 16)))
            return e;
        if (d_weights) {
            if ((e = db.get(&wm1, (size_t)n))) return e;
            if ((e = build_mask_words(p, &d_weights, 1, wm1, 0))) return e;
        }
        std::vector<double> dbh(K, 0.0), gvec(C, 0.0), hx;
        const int sweeps = 1;
        for (int sw = 0; sw < sweeps; ++sw) {
            FEHIP_CHECK(copy_sync(dbeta, beta, sizeof(double) * K, hipMemcpyHostToDevice), "beta copy");
            hipLaunchKernelGGL(combo_residual_kernel, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, 0, d_y, d_W, ld, K,
                               (const double*)dbeta, n, ec);
            if (s.S > 0)
                hipLaunchKernelGGL(combo_tables_kernel, dim3(stream_grid(s.S, kBlock)), dim3(kBlock), 0, 0,
                                   (const double*)s.alpha, s.S, K, (const double*)dbeta, ae);
            FEHIP_LAUNCH_CHECK("refinement residual");
            if ((e = within_sweep_with_tables(p, s.fe_mask, ec, ae, wm1, 0, et, 0))) return e;
            for (int j0 = 1; j0 < C; j0 += 16) {
                const int nj = std::min(16, C - j0);
                hipLaunchKernelGGL(xdot_kernel<float>, dim3(kRefineBlocks), dim3(kBlock), 0, 0, (const float*)work32,
                                   n, j0, nj, (const double*)et, xpart);
                FEHIP_LAUNCH_CHECK("refinement dots");
                std::vector<double> gs;
                if ((e = host_sum_parts(xpart, kRefineBlocks, 16, gs))) return e;
                for (int k = 0; k < nj; ++k) gvec[j0 + k] = gs[k];
            }
            std::vector<double> db_full(K, 0.0);  // (2SLS requests are stored in FP64, never here)
            for (int k = 0; k < K; ++k) db_full[k] = gvec[1 + k];
            std::fill(dbh.begin(), dbh.end(), 0.0);
            for (int a2 = 0; a2 < kept; ++a2) {
                double v = 0.0;
                for (int c2 = 0; c2 < kept; ++c2) v += Qinv[(size_t)c2 * kept + a2] * db_full[keep[c2]];
                dbh[keep[a2]] = v;
            }
            for (int k = 0; k < K; ++k) beta[k] += dbh[k];
        }
        // residual of the refined beta: e~ of the last sweep minus X~ times the last correction
        FEHIP_CHECK(copy_sync(dbeta, dbh.data(), sizeof(double) * K, hipMemcpyHostToDevice), "correction copy");
        hipLaunchKernelGGL(resid_update_kernel<float>, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, 0,
                           (const double*)et, (const float*)work32, n, K, (const double*)dbeta, r);
        FEHIP_LAUNCH_CHECK("refined residual");
    }
    if (!rss_done && (e = sumsq(r, n, part, &rss))) return e;
    result->rss = rss;
    int64_t df = std::max<int64_t>(1, N - (int64_t)(kept + dof_fes + 1));
    result->dof_resid = df;
    const double sigma2 = rss / (double)df;
    std::vector<double> cov((size_t)kept * kept, 0.0);
    bool clustered = ncd > 0;
    if (clustered) {
        // Regressors for the scores: W~, or for 2SLS X^ = Z_aug proj (fe_pipeline.f90:1181-1186).  The 2SLS
        // meat is formed without X^: the cluster sums of r X^ are proj' times those of r Z_aug, so each
        // subset's meat is proj' M_z proj with M_z the meat of the scores r Z_aug.
        const int Kx = K;
        const int K = use_iv ? n_aug : Kx;  // width of the scores (shadows the regressor count below)
        const double* const* X = f32 ? nullptr : dcols + 1;    // W~ columns
        const float* const* X32 = f32 ? dcols32 + 1 : nullptr;
        if (use_iv) X = iv_scores;  // 2SLS requests keep their own FP64 work buffer; Z_aug read in place
        // scores T = r x (row-major n x K), formed only if a subset's method reads it
        double* T = T_early;
        auto need_T = [&]() -> int {
            if (T) return 0;
            int e2 = db.get(&T, (size_t)n * K);
            if (e2) return e2;
            if (X)
                hipLaunchKernelGGL(scores_rows_kernel<double>, dim3(stream_grid((size_t)n, kBlock)), dim3(kBlock), 0,
                                   0, (const double*)r, X, n, K, T);
            else
                hipLaunchKernelGGL(scores_rows_kernel<float>, dim3(stream_grid((size_t)n, kBlock)), dim3(kBlock), 0, 0,
                                   (const double*)r, X32, n, K, T);
            FEHIP_LAUNCH_CHECK("scores rows");
            return 0;
        };
        double* S = nullptr;
        int64_t S_cap = 0;
        int64_t* dcount = nullptr;
        unsigned long long* ucount = nullptr;
        if ((e = db.get(&dcount, 1)) || (e = db.get(&ucount, 1))) return e;
        double* Md = nullptr;
        if ((e = db.get(&Md, (size_t)K * K))) return e;
        // the signed sum of the subsets' meats (the sandwich is linear in the meat, so the 2SLS
        // projection and Q^-1 M Q^-1 are applied once, to the sum)
        std::vector<double> M((size_t)K * K), Macc((size_t)K * K, 0.0);
        int64_t min_single = 0x7fffffff;
        // The fused pass over the wave tiles: the meats of the subsets it forms (a subset whose clusters are
        // single rows repeats the meat of the first such subset), and the rss of a residual formed in it.
        std::map<int, std::pair<std::vector<double>, int64_t>> pre;  // subset mask -> (meat, clusters)
        double* tsort = nullptr;  // score rows of the sorted-route subset, in its partition's order
        if (plan.fused) {
            const int KT = plan.KT;
            const int NP = KT * (KT + 1) / 2;
            WaveScores ws = plan.ws;
            ws.r = resid_in_pass ? nullptr : r;
            ws.y = resid_in_pass ? ycol : nullptr;
            ws.beta = dbeta;
            for (int k = 0; k < 12; ++k) ws.betah[k] = k < K ? beta[k] : 0.0;
            ws.K = K;
            ws.idP = p->d[p->P].ids;
            ws.wtile_rec = p->wtile_rec;
            ws.n_wtiles = p->n_wtiles;
            ws.w = wpanel;
            const size_t lds = sizeof(double) * (size_t)ws.small_lds;
            const size_t need = (size_t)((p->n_wtiles + kBlock / kWave - 1) / (kBlock / kWave));
            unsigned grid = 1;
            const int NS = ws.nsub <= 1 ? 1 : (ws.nsub <= 3 ? ws.nsub : 4);
            double *npart = nullptr, *rssp = nullptr, *spart = nullptr;
            unsigned long long* ucounts = nullptr;
            // the kernel for the tile width, the column type and the nested-subset bound; `launch` false
            // only sizes the grid (one round of resident workgroups)
            auto scores_kernel = [&](bool launch) {
#define FEHIP_SCORES_K(KTV, TXV, NSV, XV)                                                                   \
    do {                                                                                                    \
        if (!launch) {                                                                                      \
            grid = resident_grid(wave_scores_kernel<KTV, TXV, NSV>, kBlock, lds, need);                    \
        } else {                                                                                            \
            hipLaunchKernelGGL((wave_scores_kernel<KTV, TXV, NSV>), dim3(grid), dim3(kBlock), lds, 0, ws, XV, \
                               npart, ucounts);                                                             \
        }                                                                                                   \
    } while (0)
#define FEHIP_SCORES_NS(KTV, TXV, XV)        \
    if (NS == 1)                             \
        FEHIP_SCORES_K(KTV, TXV, 1, XV);     \
    else if (NS == 2)                        \
        FEHIP_SCORES_K(KTV, TXV, 2, XV);     \
    else if (NS == 3)                        \
        FEHIP_SCORES_K(KTV, TXV, 3, XV);     \
    else                                     \
        FEHIP_SCORES_K(KTV, TXV, 4, XV);
#define FEHIP_SCORES_KT(TXV, XV)             \
    if (KT == 4) {                           \
        FEHIP_SCORES_NS(4, TXV, XV)          \
    } else if (KT == 8) {                    \
        FEHIP_SCORES_NS(8, TXV, XV)          \
    } else {                                 \
        FEHIP_SCORES_NS(12, TXV, XV)         \
    }
                if (X32) {
                    FEHIP_SCORES_KT(float, X32)
                } else {
                    FEHIP_SCORES_KT(double, X)
                }
#undef FEHIP_SCORES_KT
#undef FEHIP_SCORES_NS
#undef FEHIP_SCORES_K
            };
            scores_kernel(false);
            if ((e = db.get(&npart, (size_t)std::max(1, ws.nsub) * grid * NP)) || (e = db.get(&ucounts, kMaxNestSubsets)) ||
                (e = db.get(&rssp, grid)) || (e = db.get(&spart, (size_t)grid * ws.small_lds + ws.small_lds + 1)))
                return e;
            for (int j = 0; j < ws.nrun; ++j)
                if ((e = db.get(&ws.runbuf[j], (size_t)p->d[plan.run_dim[j]].n_runs * KT + 1))) return e;
            ws.small_part = spart;
            ws.rss_part = rssp;
            int* tkey = nullptr;
            if (plan.sorted_mask) {
                if ((e = partition_positions(p, plan.sorted_dims, &ws.spos, 0))) return e;
                if ((e = db.get(&tsort, (size_t)n * K))) return e;
                ws.tsort = tsort;
                if (plan.route[plan.sorted_mask] == kRouteRunCells) {
                    if ((e = db.get(&tkey, (size_t)n))) return e;
                    ws.tkey = tkey;
                    // the scores of FP32 columns are FP32 values: the run-cells route takes them as such
                    ws.tsort32 = X32 ? 1 : 0;
                }
            }
            FEHIP_CHECK(hipMemsetAsync(ucounts, 0, sizeof(unsigned long long) * kMaxNestSubsets, 0), "memset");
            scores_kernel(true);
            FEHIP_LAUNCH_CHECK("wave scores");
            if (resid_in_pass && !f32) {  // FP64 columns: rss of the residual formed in the pass
                std::vector<double> ps;
                if ((e = host_sum_parts(rssp, (int)grid, 1, ps))) return e;
                result->rss = ps[0];
            }
            unsigned long long hc[kMaxNestSubsets] = {0, 0, 0, 0};
            if (ws.nsub > 0)
                FEHIP_CHECK(copy_sync(hc, ucounts, sizeof(hc), hipMemcpyDeviceToHost), "count copy");
            std::vector<double> hS;
            if (ws.nsmall > 0) {
                double* Sc = spart + (size_t)grid * ws.small_lds;
                hipLaunchKernelGGL(sum_blocks_kernel, dim3((unsigned)ws.small_lds), dim3(kBlock), 0, 0,
                                   (const double*)spart, (int)grid, ws.small_lds, Sc);
                FEHIP_LAUNCH_CHECK("small scores");
                hS.resize((size_t)ws.small_lds);
                FEHIP_CHECK(copy_sync(hS.data(), Sc, hS.size() * sizeof(double), hipMemcpyDeviceToHost), "S copy");
            }
            for (int mask = 1; mask < (1 << ncd); ++mask) {
                const int rt = plan.route[mask], ix = plan.index[mask];
                if (rt == kRouteNest) {
                    std::vector<double> Mq;
                    if ((e = reduce_upper(npart + (size_t)ix * grid * NP, (int)grid, KT, K, Mq))) return e;
                    pre[mask] = std::make_pair(Mq, (int64_t)hc[ix]);
                } else if (rt == kRouteRun) {
                    const DimStruct& D = p->d[plan.run_dim[ix]];
                    double* Sg = nullptr;
                    if ((e = db.get(&Sg, (size_t)D.G * KT))) return e;
                    const size_t tot = (size_t)D.G * KT;
                    if (KT == 4)
                        hipLaunchKernelGGL(group_sum_kernel<4>, dim3(stream_grid(tot, kBlock)), dim3(kBlock), 0, 0,
                                           (const double*)ws.runbuf[ix], (const int64_t*)D.grp_run_off, D.G, Sg);
                    else if (KT == 8)
                        hipLaunchKernelGGL(group_sum_kernel<8>, dim3(stream_grid(tot, kBlock)), dim3(kBlock), 0, 0,
                                           (const double*)ws.runbuf[ix], (const int64_t*)D.grp_run_off, D.G, Sg);
                    else
                        hipLaunchKernelGGL(group_sum_kernel<12>, dim3(stream_grid(tot, kBlock)), dim3(kBlock), 0, 0,
                                           (const double*)ws.runbuf[ix], (const int64_t*)D.grp_run_off, D.G, Sg);
                    FEHIP_LAUNCH_CHECK("group sums");
                    std::vector<double> Mr;
                    if ((e = meat_fused(Sg, K, KT, nullptr, nullptr, D.G, mpart, Mr))) return e;
                    pre[mask] = std::make_pair(Mr, (int64_t)Geff[plan.run_dim[ix]]);
                } else if (rt == kRouteSmall) {
                    std::vector<double> Ms((size_t)K * K, 0.0);
                    int64_t nc = 0;
                    const double* tab = hS.data() + ws.soff[ix];
                    for (int q = 0; q < plan.small_cells[ix]; ++q) {
                        const double* row = tab + (size_t)q * (KT + 1);
                        if (row[KT] > 0.0) ++nc;
                        for (int a2 = 0; a2 < K; ++a2)
                            for (int b2 = 0; b2 < K; ++b2) Ms[(size_t)b2 * K + a2] += row[a2] * row[b2];
                    }
                    pre[mask] = std::make_pair(Ms, nc);
                } else if (rt == kRouteRunCells) {
                    const Partition* pa = nullptr;
                    if ((e = panel_partition(p, plan.sorted_dims, &pa, 0))) return e;
                    FEHIP_CHECK(hipMemsetAsync(ucount, 0, sizeof(unsigned long long), 0), "memset");
                    const size_t cneed = (size_t)((pa->n_clusters + kBlock / kWave - 1) / (kBlock / kWave));
                    unsigned cgrid = 1;
#define FEHIP_RUN_CELLS_T(KTV, TTV)                                                                               \
    cgrid = std::min<unsigned>(kMeatBlocks, resident_grid(run_cells_meat_kernel<KTV, TTV>, kBlock, 0, cneed));   \
    hipLaunchKernelGGL((run_cells_meat_kernel<KTV, TTV>), dim3(cgrid), dim3(kBlock), 0, 0, (const TTV*)tsort, K, \
                       (const int*)tkey, (const int64_t*)pa->off, pa->n_clusters, plan.cells, mpart, ucount);
#define FEHIP_RUN_CELLS(KTV)                  \
    if (ws.tsort32) {                         \
        FEHIP_RUN_CELLS_T(KTV, float)         \
    } else {                                  \
        FEHIP_RUN_CELLS_T(KTV, double)        \
    }
                    if (KT == 4) {
                        FEHIP_RUN_CELLS(4)
                    } else if (KT == 8) {
                        FEHIP_RUN_CELLS(8)
                    } else {
                        FEHIP_RUN_CELLS(12)
                    }
#undef FEHIP_RUN_CELLS
#undef FEHIP_RUN_CELLS_T
                    FEHIP_LAUNCH_CHECK("run cells meat");
                    std::vector<double> Mc;
                    if ((e = reduce_upper(mpart, (int)cgrid, KT, K, Mc))) return e;
                    unsigned long long h = 0;
                    FEHIP_CHECK(copy_sync(&h, ucount, sizeof(h), hipMemcpyDeviceToHost), "count copy");
                    pre[mask] = std::make_pair(Mc, (int64_t)h);
                }
            }
            for (int mask = 1; mask < (1 << ncd); ++mask)
                if (plan.route[mask] == kRouteAlias) pre[mask] = pre[plan.index[mask]];
        }
        if (plan.wide_seg || plan.wide_run) {
            // the wide pass: segment sums and D's run sums (group order) of r x, then per subset the
            // cluster sums and their meat S'S on the matrix cores
            const int ld = (K + kWideCols - 1) / kWideCols * kWideCols;
            const int64_t GP = p->d[p->P].G;
            const DimStruct* D = plan.wide_run ? &p->d[plan.wide_dim] : nullptr;
            double *Sseg = nullptr, *Rg = nullptr;
            uint8_t* cont = nullptr;
            unsigned long long* wcounts = nullptr;
            if (plan.wide_seg && (e = db.get(&Sseg, (size_t)GP * ld))) return e;
            if (D && ((e = db.get(&Rg, (size_t)D->n_runs * ld)) || (e = db.get(&cont, (size_t)D->n_runs)))) return e;
            if ((e = db.get(&wcounts, 2))) return e;
            FEHIP_CHECK(hipMemsetAsync(wcounts, 0, 2 * sizeof(unsigned long long), 0), "memset");
            const int* Dids = D ? D->ids : nullptr;
            const int64_t* Droff = D ? D->seg_run_off : nullptr;
            const int* Ddest = D ? D->run_dest : nullptr;
            if (X32) {
                const unsigned grid = resident_grid(wide_scores_kernel<float>, kBlock, 0, (size_t)p->n_tiles);
                hipLaunchKernelGGL(wide_scores_kernel<float>, dim3(grid), dim3(kBlock), 0, 0, (const double*)r, X32, K,
                                   ld, p->seg_off, p->tile_seg, p->n_tiles, Dids, Droff, Ddest, wpanel, Sseg, Rg,
                                   cont, wcounts);
            } else {
                const unsigned grid = resident_grid(wide_scores_kernel<double>, kBlock, 0, (size_t)p->n_tiles);
                hipLaunchKernelGGL(wide_scores_kernel<double>, dim3(grid), dim3(kBlock), 0, 0, (const double*)r, X, K,
                                   ld, p->seg_off, p->tile_seg, p->n_tiles, Dids, Droff, Ddest, wpanel, Sseg, Rg,
                                   cont, wcounts);
            }
            FEHIP_LAUNCH_CHECK("wide scores");
            unsigned long long hw[2] = {0, 0};
            if (D) FEHIP_CHECK(copy_sync(hw, wcounts, sizeof(hw), hipMemcpyDeviceToHost), "count copy");
            // S'S at the padded width ld (zero columns past K, an even width for the pair loads), then
            // its leading K x K block
            double* Mdw = nullptr;
            if ((e = db.get(&Mdw, (size_t)ld * ld))) return e;
            std::vector<double> hw_meat((size_t)ld * ld);
            auto meat_rows = [&](const double* Sr, int64_t nS, std::vector<double>& Mo) -> int {
                int e2 = gemm_impl(false, true, ld, ld, nS, 1.0, Sr, ld, Sr, ld, 0.0, Mdw, ld, true);
                if (e2) return e2;
                FEHIP_CHECK(copy_sync(hw_meat.data(), Mdw, hw_meat.size() * sizeof(double), hipMemcpyDeviceToHost),
                            "meat copy");
                Mo.assign((size_t)K * K, 0.0);
                for (int j = 0; j < K; ++j)
                    for (int i = 0; i <= j; ++i) {
                        const double v = hw_meat[(size_t)j * ld + i];
                        Mo[(size_t)j * K + i] = v;
                        Mo[(size_t)i * K + j] = v;
                    }
                return 0;
            };
            bool merged = false;
            for (int mask = 1; mask < (1 << ncd); ++mask) {
                if (plan.route[mask] != kRouteWide) continue;
                bool hp = false, hd = false;
                for (int j = 0; j < ncd; ++j)
                    if (mask & (1 << j)) {
                        hp |= cdims[j] - 1 == p->P;
                        hd |= cdims[j] - 1 == plan.wide_dim;
                    }
                std::vector<double> Mw;
                int64_t nw = 0;
                if (hp && !hd) {
                    if ((e = meat_rows(Sseg, GP, Mw))) return e;
                    nw = Geff[p->P];
                } else if (hd && !hp) {  // group sums: the runs of a group are contiguous in group order
                    double* Sd = nullptr;
                    if ((e = db.get(&Sd, (size_t)D->G * ld))) return e;
                    hipLaunchKernelGGL(row_group_sums_kernel, dim3(stream_grid((size_t)D->G * ld, kBlock)),
                                       dim3(kBlock), 0, 0, (const double*)Rg, ld, (const int64_t*)D->grp_run_off,
                                       (int64_t)D->G, Sd);
                    FEHIP_LAUNCH_CHECK("group sums");
                    if ((e = meat_rows(Sd, D->G, Mw))) return e;
                    db.release(Sd);
                    nw = Geff[plan.wide_dim];
                } else {  // the cells (segment, group): runs, merged where a segment returns to a group
                    if (hw[1] > 0 && !merged) {
                        hipLaunchKernelGGL(merge_cells_kernel, dim3(stream_grid((size_t)D->n_runs * ld, kBlock)),
                                           dim3(kBlock), 0, 0, Rg, ld, (const uint8_t*)cont, D->n_runs);
                        FEHIP_LAUNCH_CHECK("merge cells");
                        merged = true;
                    }
                    if ((e = meat_rows(Rg, D->n_runs, Mw))) return e;
                    nw = (int64_t)hw[0];
                }
                pre[mask] = std::make_pair(Mw, nw);
            }
        }
        for (int mask = 1; mask < (1 << ncd); ++mask) {
            unsigned dims_mask = 0;
            int len = 0;
            for (int j = 0; j < ncd; ++j)
                if (mask & (1 << j)) {
                    dims_mask |= 1u << (cdims[j] - 1);
                    ++len;
                }
            const double wgt = (len % 2 == 1) ? 1.0 : -1.0;
            // route the subset to the cheapest exact method (see the kernels above)
            bool has_p = (dims_mask >> p->P) & 1u;
            bool all_small = true;
            int64_t cells = 1;
            for (int d = 0; d < p->m; ++d)
                if ((dims_mask >> d) & 1u) {
                    all_small &= p->d[d].small != 0;
                    cells *= p->d[d].G;
                }
            const int single = len == 1 ? __builtin_ctz(dims_mask) : -1;
            int64_t ncl = 0;
            // the sorted-route subset reads the score rows the fused pass wrote in its partition's order
            const bool in_order = plan.route[mask] == kRouteSorted;
            if (pre.count(mask)) {
                M = pre[mask].first;
                ncl = pre[mask].second;
            } else if (!in_order && has_p && p->n_long == 0 && K <= 12) {
                if ((e = need_T())) return e;
                NestKeys nk{};
                for (int d = 0; d < p->m; ++d)
                    if (d != p->P && ((dims_mask >> d) & 1u)) nk.ids[nk.nk++] = p->d[d].ids;
                FEHIP_CHECK(hipMemsetAsync(ucount, 0, sizeof(unsigned long long), 0), "memset");
                const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(p->n_tiles, kMeatBlocks));
                const int KT = K <= 4 ? 4 : (K <= 8 ? 8 : 12);
                if (KT == 4)
                    hipLaunchKernelGGL(nested_meat_kernel<4>, dim3(grid), dim3(kBlock), 0, 0, T, K, p->seg_off,
                                       p->tile_seg, p->n_tiles, nk, wpanel, mpart, ucount);
                else if (KT == 8)
                    hipLaunchKernelGGL(nested_meat_kernel<8>, dim3(grid), dim3(kBlock), 0, 0, T, K, p->seg_off,
                                       p->tile_seg, p->n_tiles, nk, wpanel, mpart, ucount);
                else
                    hipLaunchKernelGGL(nested_meat_kernel<12>, dim3(grid), dim3(kBlock), 0, 0, T, K, p->seg_off,
                                       p->tile_seg, p->n_tiles, nk, wpanel, mpart, ucount);
                FEHIP_LAUNCH_CHECK("nested meat");
                if ((e = reduce_upper(mpart, (int)grid, KT, K, M))) return e;
                unsigned long long h = 0;
                FEHIP_CHECK(copy_sync(&h, ucount, sizeof(h), hipMemcpyDeviceToHost), "count copy");
                ncl = (int64_t)h;
            } else if (!in_order && all_small && cells <= kSmallCells && K <= 11) {
                SmallKeys sk{};
                int stride = 1;
                for (int d = 0; d < p->m; ++d)
                    if ((dims_mask >> d) & 1u) {
                        sk.ids[sk.nk] = p->d[d].ids;
                        sk.stride[sk.nk++] = stride;
                        stride *= p->d[d].G;
                    }
                sk.w = wpanel;
                const int KT = K <= 3 ? 4 : (K <= 7 ? 8 : 12);
                const int c = (int)cells;
                // sixteen waves per CU, fewer where the per-block partial tables would pass 4M doubles
                const unsigned grid = std::max(
                    1u, std::min(stream_grid((size_t)n, kBlock, 16), (unsigned)std::max<int64_t>(
                                                                          256, ((int64_t)4 << 20) / ((int64_t)c * KT))));
                double* sp_ = nullptr;
                if ((e = db.get(&sp_, (size_t)grid * c * KT + (size_t)c * KT))) return e;
                const size_t lds = sizeof(double) * (size_t)c * KT;
#define FEHIP_SMALL(KTV)                                                                                        \
    if (X32)                                                                                                    \
        hipLaunchKernelGGL((small_scores_kernel<KTV, float>), dim3(grid), dim3(kBlock), lds, 0, (const double*)T,  \
                           (const double*)r, X32, K, n, sk, c, sp_);                                            \
    else                                                                                                        \
        hipLaunchKernelGGL((small_scores_kernel<KTV, double>), dim3(grid), dim3(kBlock), lds, 0, (const double*)T, \
                           (const double*)r, X, K, n, sk, c, sp_);
                if (KT == 4) {
                    FEHIP_SMALL(4)
                } else if (KT == 8) {
                    FEHIP_SMALL(8)
                } else {
                    FEHIP_SMALL(12)
                }
#undef FEHIP_SMALL
                double* Sc = sp_ + (size_t)grid * c * KT;
                hipLaunchKernelGGL(sum_blocks_kernel, dim3((unsigned)(c * KT)), dim3(kBlock), 0, 0, sp_, (int)grid,
                                   c * KT, Sc);
                FEHIP_LAUNCH_CHECK("small scores");
                std::vector<double> hS((size_t)c * KT);
                FEHIP_CHECK(copy_sync(hS.data(), Sc, hS.size() * sizeof(double), hipMemcpyDeviceToHost), "S copy");
                M.assign((size_t)K * K, 0.0);
                for (int q = 0; q < c; ++q) {
                    if (hS[(size_t)q * KT + KT - 1] > 0.0) ++ncl;
                    for (int a2 = 0; a2 < K; ++a2)
                        for (int b2 = 0; b2 < K; ++b2)
                            M[(size_t)b2 * K + a2] += hS[(size_t)q * KT + a2] * hS[(size_t)q * KT + b2];
                }
            } else if (!in_order && single >= 0 && single != p->P && !p->d[single].small && p->n_long == 0 &&
                       K <= 12) {
                const DimStruct& D = p->d[single];
                const int KT = K <= 4 ? 4 : (K <= 8 ? 8 : 12);
                double *rb = nullptr, *Sg = nullptr;
                if ((e = db.get(&rb, (size_t)D.n_runs * KT + 1)) || (e = db.get(&Sg, (size_t)D.G * KT))) return e;
                const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(p->n_tiles, (int64_t)num_cus() * 8));
#define FEHIP_RUNS(KTV)                                                                                          \
    if (X32)                                                                                                     \
        hipLaunchKernelGGL((run_scores_kernel<KTV, float>), dim3(grid), dim3(kBlock), 0, 0, (const double*)T,       \
                           (const double*)r, X32, K, p->seg_off, p->tile_seg, p->n_tiles, D.ids, D.seg_run_off, rb); \
    else                                                                                                         \
        hipLaunchKernelGGL((run_scores_kernel<KTV, double>), dim3(grid), dim3(kBlock), 0, 0, (const double*)T,      \
                           (const double*)r, X, K, p->seg_off, p->tile_seg, p->n_tiles, D.ids, D.seg_run_off, rb);
                if (KT == 4) {
                    FEHIP_RUNS(4)
                } else if (KT == 8) {
                    FEHIP_RUNS(8)
                } else {
                    FEHIP_RUNS(12)
                }
#undef FEHIP_RUNS
                FEHIP_LAUNCH_CHECK("run scores");
                const size_t tot = (size_t)D.G * KT;
                if (KT == 4)
                    hipLaunchKernelGGL(group_runs_kernel<4>, dim3(stream_grid(tot, kBlock)), dim3(kBlock), 0, 0, rb,
                                       D.run_order, D.grp_run_off, D.G, Sg);
                else if (KT == 8)
                    hipLaunchKernelGGL(group_runs_kernel<8>, dim3(stream_grid(tot, kBlock)), dim3(kBlock), 0, 0, rb,
                                       D.run_order, D.grp_run_off, D.G, Sg);
                else
                    hipLaunchKernelGGL(group_runs_kernel<12>, dim3(stream_grid(tot, kBlock)), dim3(kBlock), 0, 0, rb,
                                       D.run_order, D.grp_run_off, D.G, Sg);
                FEHIP_LAUNCH_CHECK("group runs");
                if ((e = meat_fused(Sg, K, KT, nullptr, nullptr, D.G, mpart, M))) return e;
                ncl = Geff[single];
            } else {
                if (!in_order && (e = need_T())) return e;
                const double* Tm = in_order ? tsort : T;
                const Partition* part_s = nullptr;
                if ((e = panel_partition(p, dims_mask, &part_s, 0))) return e;
                ncl = part_s->n_clusters;
                const int64_t nS = part_s->n_clusters;
                // a partition of singletons needs no row order: sum_i T_i T_i'
                const bool singletons = nS == n;
                const int* ord = (singletons || in_order) ? nullptr : part_s->order;
                const int64_t* off = singletons ? nullptr : part_s->off;
                if (K <= 12 && singletons) {
                    if ((e = meat_fused(Tm, K, K, nullptr, nullptr, n, mpart, M))) return e;
                } else if (K <= 12) {
                    const int KT = K <= 4 ? 4 : (K <= 8 ? 8 : 12);
                    const int64_t nwaves = (n + kWave - 1) / kWave;
                    int* bkey = nullptr;
                    double *bval = nullptr, *Sb = nullptr;
                    if ((e = db.get(&bkey, (size_t)(2 * nwaves))) || (e = db.get(&bval, (size_t)(2 * nwaves * KT))) ||
                        (e = db.get(&Sb, (size_t)(2 * nwaves * KT))))
                        return e;
                    const unsigned grid = (unsigned)std::max<int64_t>(
                        1, std::min<int64_t>((nwaves + kBlock / kWave - 1) / (kBlock / kWave), (int64_t)kMeatBlocks));
                    if (KT == 4)
                        hipLaunchKernelGGL(sorted_wave_meat_kernel<4>, dim3(grid), dim3(kBlock), 0, 0, (const double*)Tm,
                                           K, ord, (const int*)part_s->keys, n, mpart, bkey, bval);
                    else if (KT == 8)
                        hipLaunchKernelGGL(sorted_wave_meat_kernel<8>, dim3(grid), dim3(kBlock), 0, 0, (const double*)Tm,
                                           K, ord, (const int*)part_s->keys, n, mpart, bkey, bval);
                    else
                        hipLaunchKernelGGL(sorted_wave_meat_kernel<12>, dim3(grid), dim3(kBlock), 0, 0,
                                           (const double*)Tm, K, ord, (const int*)part_s->keys, n,
                                           mpart, bkey, bval);
                    FEHIP_LAUNCH_CHECK("sorted wave meat");
                    if ((e = reduce_upper(mpart, (int)grid, KT, K, M))) return e;
                    // clusters cut by wave boundaries: merge their pieces, then add their S S'
                    if (KT == 4)
                        e = cluster_sums<4>(bval, KT, nullptr, bkey, 2 * nwaves, Sb, dcount, 0);
                    else if (KT == 8)
                        e = cluster_sums<8>(bval, KT, nullptr, bkey, 2 * nwaves, Sb, dcount, 0);
                    else
                        e = cluster_sums<12>(bval, KT, nullptr, bkey, 2 * nwaves, Sb, dcount, 0);
                    if (e) return e;
                    int64_t n_groups = 0;
                    FEHIP_CHECK(copy_sync(&n_groups, dcount, sizeof(int64_t), hipMemcpyDeviceToHost), "count copy");
                    std::vector<double> Mb;
                    if ((e = meat_fused(Sb, K, KT, nullptr, nullptr, n_groups, mpart, Mb))) return e;
                    for (size_t i2 = 0; i2 < M.size(); ++i2) M[i2] += Mb[i2];
                } else {
                    if (nS * K > S_cap) {
                        if (S) db.release(S);
                        S_cap = nS * K;
                        if ((e = db.get(&S, (size_t)S_cap))) return e;
                    }
                    hipLaunchKernelGGL(segsum_kernel, dim3(stream_grid((size_t)nS * K, kBlock)), dim3(kBlock), 0, 0,
                                       (const double*)Tm, K, ord, off, nS, S);
                    FEHIP_LAUNCH_CHECK("segsum");
                    if ((e = gemm_impl(false, true, K, K, nS, 1.0, S, K, S, K, 0.0, Md, K, true))) return e;
                    FEHIP_CHECK(copy_sync(M.data(), Md, M.size() * sizeof(double), hipMemcpyDeviceToHost), "meat copy");
                    for (int j = 0; j < K; ++j)
                        for (int i = j + 1; i < K; ++i) M[(size_t)j * K + i] = M[(size_t)i * K + j];
                }
                if (wpanel) {  // count the clusters that hold a weighted row
                    int* mark = nullptr;
                    if ((e = db.get(&mark, (size_t)nS + 2))) return e;
                    FEHIP_CHECK(hipMemsetAsync(mark, 0, sizeof(int) * ((size_t)nS + 2), 0), "memset");
                    FEHIP_CHECK(hipMemsetAsync(ucount, 0, sizeof(unsigned long long), 0), "memset");
                    hipLaunchKernelGGL(mark_kernel, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, 0, part_s->keys,
                                       part_s->order, wpanel, n, mark);
                    hipLaunchKernelGGL(count_kernel, dim3(stream_grid((size_t)nS + 2, kBlock)), dim3(kBlock), 0, 0,
                                       mark, (int64_t)nS + 2, ucount);
                    unsigned long long h = 0;
                    FEHIP_CHECK(copy_sync(&h, ucount, sizeof(h), hipMemcpyDeviceToHost), "count copy");
                    ncl = (int64_t)h;
                }
            }
            if (len == 1) {
                ncl = Geff[single];
                min_single = std::min<int64_t>(min_single, ncl);
            }
            result->n_clusters[mask - 1] = ncl;
            for (size_t i2 = 0; i2 < M.size(); ++i2) Macc[i2] += wgt * M[i2];
        }
        // 2SLS: the meat of r X^ is proj' M_z proj
        const std::vector<double>* Mp = &Macc;
        std::vector<double> Mx;
        if (use_iv && (e = project_meat(Macc, n_aug, proj, Kx, Mx))) return e;
        if (use_iv) Mp = &Mx;
        // cov = Qinv M_kept Qinv (column-major; the inner loops run down contiguous columns)
        {
            std::vector<double> Mk((size_t)kept * kept), tmp((size_t)kept * kept, 0.0);
            for (int c = 0; c < kept; ++c)
                for (int a = 0; a < kept; ++a) Mk[(size_t)c * kept + a] = (*Mp)[(size_t)keep[c] * Kx + keep[a]];
            for (int c = 0; c < kept; ++c)
                for (int k = 0; k < kept; ++k) {
                    const double q = Qinv[(size_t)c * kept + k];
                    const double* mk = Mk.data() + (size_t)k * kept;
                    double* tc = tmp.data() + (size_t)c * kept;
                    for (int a = 0; a < kept; ++a) tc[a] += mk[a] * q;
                }
            for (int c = 0; c < kept; ++c)
                for (int k = 0; k < kept; ++k) {
                    const double t = tmp[(size_t)c * kept + k];
                    const double* qk = Qinv.data() + (size_t)k * kept;
                    double* cc = cov.data() + (size_t)c * kept;
                    for (int a = 0; a < kept; ++a) cc[a] += qk[a] * t;
                }
        }
        const int nested_adj = dof_nested > 0 ? 1 : 0;
        const int64_t den = std::max<int64_t>(1, N - (kept + nested_adj));
        double scalar = (double)std::max<int64_t>(1, N - 1) / (double)den;
        if (min_single > 1) scalar *= (double)min_single / (double)(min_single - 1);
        for (double& v : cov) v *= scalar;
        for (int j = 0; j < kept; ++j)
            for (int i = j + 1; i < kept; ++i) cov[(size_t)j * kept + i] = cov[(size_t)i * kept + j];
        result->psd_fix_applied = dense::psd_clip(cov, kept);
        result->min_single_cluster = min_single;
        result->dof_resid = std::max<int64_t>(1, min_single - 1);
    } else {
        for (int a = 0; a < kept; ++a)
            for (int c = 0; c < kept; ++c) cov[(size_t)c * kept + a] = sigma2 * Qinv[(size_t)c * kept + a];
    }
    for (int a = 0; a < kept; ++a) se[keep[a]] = std::sqrt(std::max(0.0, cov[(size_t)a * kept + a]));
    if (vcov)
        for (int a = 0; a < kept; ++a)
            for (int c = 0; c < kept; ++c) vcov[(size_t)keep[c] * K + keep[a]] = cov[(size_t)c * kept + a];
    finalize_stats(result, N);
    FEHIP_CHECK(hipStreamSynchronize(0), "estimate sync");
    result->t_se = std::chrono::duration<double>(clk::now() - t2).count();
    return 0;
}

constexpr int kMaxSlots = 32;

// a chunk whose distinct columns number at most this many takes its cross products from the final
// pass of the within-transform (one column tile)
constexpr int kGramCols = 16;

// One chunk of a batch: requests that absorb the same FE dims and share the within-transform sweeps (up
// to 32 requests per solve, one mask bit each), and what the solve leaves for the requests' normal
// equations and sandwiches.
struct ColDst {
    double* out;
    float* out32;
    double* alpha;
};

struct ChunkRun {
    std::vector<int> idx;   // request indices
    double bytes = 0.0;     // work storage of its requests
    bool share = false;     // one copy per distinct column, shared by the requests
    bool weighted = false;  // a request has a row subsample
    unsigned mask = 0;
    int group = -1;         // group of chunks whose right-hand sides and final pass are shared
    DevBuf cb;              // the chunk's device buffers
    uint32_t* wm = nullptr;
    int ncols = 0, nu = 0;
    std::vector<ColDst> dst;
    std::vector<int> uidx, upos;
    WithinCols wc{0};
    double* dgram = nullptr;
    std::vector<double> gh;  // cross products of the solved columns (nu x nu, upper triangle at [hi * nu + lo])
    WithinStats wsu;
};

// Chunks of the same primary treatment (all absorb the primary dim, or none does), unweighted and with
// shared column storage: their right-hand sides come from one pass over the union of their FE dims and
// columns (each FE dim's block of D_o' M_P x, or D_o' x, does not depend on the other dims of the set).
struct ChunkGroup {
    unsigned mask = 0;
    std::vector<int> members;
    std::vector<const double*> xs;  // distinct input columns
    bool f32 = true;                // every request of the group stores FP32 (the pass may read FP32 copies)
    SharedRhs rhs{};
    bool has_rhs = false;
    DevBuf buf;
};

const double* req_col(const ReqState& q, int j) {
    return j == 0 ? q.d_y : (j <= q.K ? q.d_W + (int64_t)(j - 1) * q.ld : q.d_Z + (int64_t)(j - 1 - q.K) * q.ld);
}

// Preparation of the chunk's requests, its column list (a column that several requests share -- same
// input, same row subsample, same storage precision -- is solved once) and its solve.
int solve_chunk(Panel* p, std::vector<ReqState>& rs, ChunkRun& ch, const ChunkGroup* grp, BatchShared& shared) {
    int e;
    const int64_t n = p->n;
    for (int i : ch.idx) {
        rs[i].shared = ch.share;
        if ((e = prepare_request(p, rs[i], shared))) return e;
    }
    ch.ncols = 0;
    for (int i : ch.idx) ch.ncols += rs[i].C;
    if (ch.weighted) {
        std::vector<const uint8_t*> w(ch.idx.size());
        for (size_t b = 0; b < ch.idx.size(); ++b) w[b] = rs[ch.idx[b]].d_weights;
        if ((e = ch.cb.get(&ch.wm, (size_t)n))) return e;
        if ((e = build_mask_words(p, w.data(), (int)ch.idx.size(), ch.wm, 0))) return e;
    }
    const int ncols = ch.ncols;
    ch.dst.assign(ncols, ColDst{nullptr, nullptr, nullptr});
    ch.uidx.assign(ncols, 0);
    std::map<std::tuple<const double*, const uint8_t*, bool>, int> seen;
    ch.wc = WithinCols(0);
    WithinCols& wc = ch.wc;
    int c = 0;
    for (size_t b = 0; b < ch.idx.size(); ++b) {
        ReqState& s = rs[ch.idx[b]];
        if (ch.share) s.shared_cols.assign(s.C, nullptr);
        for (int j = 0; j < s.C; ++j, ++c) {
            const double* x = req_col(s, j);
            const auto key = std::make_tuple(x, s.d_weights, s.f32);
            auto it = seen.find(key);
            if (ch.share) {
                if (it == seen.end()) {  // first use: the column's one copy
                    if (s.f32 ? (e = ch.cb.get(&ch.dst[c].out32, (size_t)n)) : (e = ch.cb.get(&ch.dst[c].out, (size_t)n)))
                        return e;
                }
                const ColDst& d = ch.dst[it == seen.end() ? c : it->second];
                s.shared_cols[j] = s.f32 ? (const void*)d.out32 : (const void*)d.out;
            } else {
                ch.dst[c] = {s.work ? s.work + (int64_t)j * n : nullptr, s.work32 ? s.work32 + (int64_t)j * n : nullptr,
                             s.alpha ? s.alpha + (int64_t)j * s.S : nullptr};
            }
            if (it != seen.end() && (ch.share || (ch.dst[it->second].alpha != nullptr) == (ch.dst[c].alpha != nullptr))) {
                ch.uidx[c] = it->second;
                continue;
            }
            seen[key] = c;
            ch.uidx[c] = c;
            wc.x.push_back(x);
            wc.x32.push_back(s.f32 ? panel_col32(p, x) : nullptr);  // FP32 storage reads a registered FP32 copy
            wc.out.push_back(ch.dst[c].out);
            wc.out32.push_back(ch.dst[c].out32);
            wc.alpha.push_back(ch.dst[c].alpha);
            wc.slot.push_back(ch.wm ? (int)b : 0);
        }
    }
    ch.nu = (int)wc.x.size();
    ch.upos.assign(ncols, -1);  // position of each solved column in wc
    for (int c2 = 0, u = 0; c2 < ncols; ++c2)
        if (ch.uidx[c2] == c2) ch.upos[c2] = u++;
    const ReqState& lead = rs[ch.idx[0]];
    fehip_demean_opts dopt = lead.opts->demean;
    dopt.fe_mask = lead.fe_mask;
    if (grp && grp->has_rhs) {  // right-hand sides from the group's shared pass
        wc.rhs = &grp->rhs;
        wc.rhs_col.resize(ch.nu);
        for (int u = 0; u < ch.nu; ++u) {
            const auto f = std::find(grp->xs.begin(), grp->xs.end(), wc.x[u]);
            if (f == grp->xs.end()) return set_error("batch: a column is missing from the shared right-hand sides");
            wc.rhs_col[u] = (int)(f - grp->xs.begin());
        }
    }
    if (ch.nu <= kGramCols) {
        // cross products of the solved columns from the final pass (one column tile when they fit)
        if ((e = ch.cb.get(&ch.dgram, (size_t)ch.nu * ch.nu))) return e;
        wc.gram = ch.dgram;
        wc.gram_ld = ch.nu;
    }
    ch.wsu = WithinStats(ch.nu);
    if ((e = within_transform(p, wc, ch.nu, ch.wm, ch.wm ? (int)ch.idx.size() : 1, dopt, 1, ch.wsu, 0))) return e;
    if (ch.dgram) {
        ch.gh.assign((size_t)ch.nu * ch.nu, 0.0);
        FEHIP_CHECK(copy_sync(ch.gh.data(), ch.dgram, ch.gh.size() * sizeof(double), hipMemcpyDeviceToHost), "gram copy");
    }
    return 0;
}

// The requests' cross products, the copies of columns shared by requests with their own storage, the
// solve statistics, then every request's normal equations and sandwich.
int finish_chunk(Panel* p, std::vector<ReqState>& rs, ChunkRun& ch) {
    int e;
    const int64_t n = p->n;
    if (!ch.gh.empty()) {
        int cc = 0;
        for (int i : ch.idx) {
            ReqState& s = rs[i];
            std::vector<int> u(s.C);
            bool one_tile = true;
            for (int j = 0; j < s.C; ++j) {
                u[j] = ch.upos[ch.uidx[cc + j]];
                one_tile = one_tile && ch.wsu.col_tile[u[j]] == ch.wsu.col_tile[u[0]];
            }
            cc += s.C;
            s.gram.clear();
            if (!one_tile) continue;
            s.gram.resize((size_t)s.C * s.C);
            for (int b2 = 0; b2 < s.C; ++b2)
                for (int a2 = 0; a2 < s.C; ++a2) {  // upper triangle of the solve's matrix, mirrored
                    const int lo = std::min(u[a2], u[b2]), hi = std::max(u[a2], u[b2]);
                    s.gram[(size_t)b2 * s.C + a2] = ch.gh[(size_t)hi * ch.nu + lo];
                }
        }
    }
    const int ncols = ch.ncols;
    WithinStats ws(ncols);
    for (int c2 = 0; c2 < ncols; ++c2) {
        const int src = ch.uidx[c2], u = ch.upos[src];
        ws.col_iters[c2] = ch.wsu.col_iters[u];
        ws.col_converged[c2] = ch.wsu.col_converged[u];
        ws.col_max_update[c2] = ch.wsu.col_max_update[u];
        ws.col_refines[c2] = ch.wsu.col_refines[u];
        if (src == c2 || ch.share) continue;
        if (ch.dst[c2].out)
            FEHIP_CHECK(hipMemcpyAsync(ch.dst[c2].out, ch.dst[src].out, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, 0),
                        "shared column");
        if (ch.dst[c2].out32)
            FEHIP_CHECK(hipMemcpyAsync(ch.dst[c2].out32, ch.dst[src].out32, sizeof(float) * (size_t)n,
                                       hipMemcpyDeviceToDevice, 0),
                        "shared column");
        if (ch.dst[c2].alpha && ch.dst[src].alpha)
            FEHIP_CHECK(hipMemcpyAsync(ch.dst[c2].alpha, ch.dst[src].alpha, sizeof(double) * (size_t)within_table_size(p, ch.mask),
                                       hipMemcpyDeviceToDevice, 0),
                        "shared tables");
    }
    const auto t1 = clk::now();
    int c = 0;
    for (int i : ch.idx) {
        ReqState& s = rs[i];
        s.result->converged = 1;
        for (int j = 0; j < s.C; ++j, ++c) {
            s.result->fe_iterations = std::max(s.result->fe_iterations, ws.col_iters[c]);
            s.result->converged = s.result->converged && ws.col_converged[c];
            s.result->max_update = std::max(s.result->max_update, ws.col_max_update[c]);
            s.result->fe_refinements = std::max(s.result->fe_refinements, ws.col_refines[c]);
        }
        s.t1 = t1;
        s.result->t_demean = std::chrono::duration<double>(t1 - s.t0).count();
    }
    for (int i : ch.idx) {
        if ((e = finish_request(p, rs[i]))) return e;
        FEHIP_CHECK(hipStreamSynchronize(0), "request done");  // its kernels have read its buffers
        rs[i].db.clear();  // release the request's device buffers before the next chunk
    }
    ch.cb.clear();
    return 0;
}

// Two requests of a batch that ask for the same regression (same columns, row subsample and options) get
// the same answer: the later one is served from the earlier one's computation.
bool same_request(const ReqState& a, const ReqState& b) {
    if (a.d_y != b.d_y || a.d_W != b.d_W || a.d_Z != b.d_Z || a.K != b.K || a.L != b.L || a.ld != b.ld ||
        a.d_weights != b.d_weights || a.fe_mask != b.fe_mask)
        return false;
    const fehip_opts &x = *a.opts, &y = *b.opts;
    const fehip_demean_opts &dx = x.demean, &dy = y.demean;
    if (dx.tol != dy.tol || dx.max_iter != dy.max_iter || dx.method != dy.method || dx.precision != dy.precision ||
        dx.check_every != dy.check_every || x.n_cluster != y.n_cluster || x.n_iv_reg != y.n_iv_reg ||
        x.n_iv_inst != y.n_iv_inst)
        return false;
    for (int i = 0; i < x.n_cluster && i < 4; ++i)
        if (x.cluster_dims[i] != y.cluster_dims[i]) return false;
    for (int i = 0; i < x.n_iv_reg; ++i)
        if (x.iv_reg[i] != y.iv_reg[i]) return false;
    for (int i = 0; i < x.n_iv_inst; ++i)
        if (x.iv_inst[i] != y.iv_inst[i]) return false;
    return true;
}

// Solve a batch: requests that absorb the same FE dims share the within-transform sweeps (chunks), chunks
// of the same primary treatment share their right-hand-side pass, and every request then takes its own
// normal-equation and sandwich step (a request equal to an earlier one takes that one's results).
int run_batch(Panel* p, std::vector<ReqState>& rs) {
    int e;
    const int64_t n = p->n;
    const unsigned all = (1u << p->m) - 1u;
    for (auto& s : rs) {
        if (!s.opts) return set_error("fehip_estimate: null options");
        s.fe_mask = (s.opts->demean.fe_mask ? s.opts->demean.fe_mask : all) & all;
    }
    // device memory a chunk may take for its work buffers
    size_t free_b = 0, total_b = 0;
    FEHIP_CHECK(hipMemGetInfo(&free_b, &total_b), "memory info");
    const double budget = 0.7 * (double)(free_b + cache_idle_bytes());
    BatchShared shared;
    std::vector<int> copy_of(rs.size(), -1);
    for (size_t i = 1; i < rs.size(); ++i)
        for (size_t j = 0; j < i && copy_of[i] < 0; ++j)
            if (copy_of[j] < 0 && (!rs[i].vcov || rs[j].vcov) && same_request(rs[i], rs[j])) copy_of[i] = (int)j;
    // chunks: unsolved requests with the same FE set, in submission order, within the slot count and the
    // memory budget (the first request always goes)
    std::vector<std::unique_ptr<ChunkRun>> runs;
    const char* ms_env = getenv("FEHIP_MULTI_SLOT");
    const bool multi_slot = ms_env && ms_env[0] == '1';
    {
        std::vector<int> taken(rs.size(), 0);
        for (size_t i = 0; i < rs.size(); ++i) taken[i] = copy_of[i] >= 0;
        for (size_t i0 = 0; i0 < rs.size(); ++i0) {
            if (taken[i0]) continue;
            auto ch = std::make_unique<ChunkRun>();
            const fehip_demean_opts& lead = rs[i0].opts->demean;
            for (size_t i = i0; i < rs.size() && (int)ch->idx.size() < kMaxSlots; ++i) {
                const fehip_demean_opts& o = rs[i].opts->demean;
                if (taken[i] || rs[i].fe_mask != rs[i0].fe_mask || o.tol != lead.tol || o.max_iter != lead.max_iter ||
                    o.check_every != lead.check_every)
                    continue;
                // requests with different row subsamples share no column, right-hand side or table: one
                // multi-slot solve would iterate as long as its slowest subsample and run the multi-slot row
                // kernels, so each subsample takes its own chunk (64 requests of the 10M-row stream: 344 ms
                // one subsample per chunk, 375 ms shared); FEHIP_MULTI_SLOT=1 shares the sweeps (comparisons)
                if (rs[i].d_weights != rs[i0].d_weights && !multi_slot) continue;
                const bool f32 = o.precision == 1 && !(rs[i].opts->n_iv_reg > 0 && rs[i].L > 0);
                const double b = (f32 ? 4.0 : 8.0) * (double)n * (1 + std::max(0, rs[i].K) + std::max(0, rs[i].L)) +
                                 2.0 * (double)n;
                if (!ch->idx.empty() && ch->bytes + b > budget) break;
                ch->bytes += b;
                ch->idx.push_back((int)i);
            }
            for (int i : ch->idx) taken[i] = 1;
            ch->mask = rs[i0].fe_mask;
            // Shared storage: when the chunk's distinct columns fit one column tile and no request is a 2SLS
            // one, every request reads its transformed columns (and takes its cross products) from one copy
            // per distinct column, and no request holds a work buffer of its own.
            std::set<std::tuple<const double*, const uint8_t*, bool>> keys;
            bool share = true;
            for (int i : ch->idx) {
                const ReqState& q = rs[i];
                const bool iv = q.opts->n_iv_reg > 0 && q.L > 0;
                share = share && !iv;
                const bool f32 = q.opts->demean.precision == 1 && !iv;
                ch->weighted |= q.d_weights != nullptr;
                for (int j = 0; j < 1 + q.K + q.L; ++j) keys.insert(std::make_tuple(req_col(q, j), q.d_weights, f32));
            }
            ch->share = share && (int)keys.size() <= std::min(kGramCols, within_tile_cap(p, ch->mask));
            runs.push_back(std::move(ch));
        }
    }
    // groups of chunks that share their right-hand-side and final passes
    std::vector<std::unique_ptr<ChunkGroup>> groups;
    for (unsigned elim = 0; elim < 2; ++elim) {
        auto g = std::make_unique<ChunkGroup>();
        for (size_t ci = 0; ci < runs.size(); ++ci) {
            const ChunkRun& ch = *runs[ci];
            if (!ch.share || ch.weighted || ((ch.mask >> p->P) & 1u) != elim) continue;
            std::vector<const double*> xs = g->xs;
            for (int i : ch.idx)  // (the requests are not prepared yet: C = 1 + K + L)
                for (int j = 0; j < 1 + rs[i].K + rs[i].L; ++j)
                    if (std::find(xs.begin(), xs.end(), req_col(rs[i], j)) == xs.end()) xs.push_back(req_col(rs[i], j));
            const unsigned um = g->mask | ch.mask;
            const int cap = std::min(kGramCols, within_tile_cap(p, um));
            if ((int)xs.size() > cap) continue;
            g->xs = xs;
            g->mask = um;
            g->members.push_back((int)ci);
            for (int i : ch.idx)
                g->f32 = g->f32 && rs[i].opts->demean.precision == 1 && !(rs[i].opts->n_iv_reg > 0 && rs[i].L > 0);
        }
        if (g->members.size() < 2) continue;
        const int gi = (int)groups.size();
        for (int ci : g->members) runs[ci]->group = gi;
        groups.push_back(std::move(g));
    }
    for (auto& g : groups) {
        if (within_table_size(p, g->mask) <= 0) continue;
        double* b = nullptr;
        if ((e = g->buf.get(&b, (size_t)within_table_size(p, g->mask) * kGramCols + 16))) return e;
        int ld = 0;
        std::vector<const float*> xs32;
        if (g->f32)
            for (const double* x : g->xs) xs32.push_back(panel_col32(p, x));
        if ((e = within_rhs(p, g->mask, g->xs.data(), xs32.empty() ? nullptr : xs32.data(), (int)g->xs.size(), b, &ld,
                            0)))
            return e;
        g->rhs.b = b;
        g->rhs.ld = ld;
        if ((e = within_dim_offsets(p, g->mask, g->rhs.dim_off))) return e;
        g->has_rhs = true;
    }
    // Solve the chunks in turn; each chunk's finish (normal equations, residuals, sandwiches) runs on one helper
    // thread, on that thread's default stream, while the next chunk is solved, so the host work between the
    // sandwiches' kernels and the solves' small kernels overlap on the device.  A chunk's solve has completed on
    // the device when it returns (its cross products are read back), and a finish releases its requests' buffers
    // only after its own stream has drained; the solver waits when the helper is a chunk behind, which bounds
    // the buffers held.  While the helper runs, the solves launch their CG steps directly: the runtime refuses
    // blocking calls from any thread while a graph is being captured.
    if (runs.size() == 1) {
        ChunkRun& ch = *runs[0];
        if ((e = solve_chunk(p, rs, ch, ch.group >= 0 ? groups[ch.group].get() : nullptr, shared))) return e;
        if ((e = finish_chunk(p, rs, ch))) return e;
    } else {
        const bool capture = thread_capture();
        set_thread_capture(false);
        int device = 0;
        (void)hipGetDevice(&device);
        std::mutex qm;
        std::condition_variable qcv;
        std::deque<ChunkRun*> queue;
        int pending = 0;  // chunks handed over and not finished
        bool closed = false;
        int fin_e = 0;
        std::string fin_msg;
        std::thread helper([&]() {
            (void)hipSetDevice(device);
            set_thread_capture(false);
            for (;;) {
                ChunkRun* c = nullptr;
                {
                    std::unique_lock<std::mutex> lk(qm);
                    qcv.wait(lk, [&] { return !queue.empty() || closed; });
                    if (queue.empty()) break;
                    c = queue.front();
                    queue.pop_front();
                }
                int fe = 0;
                std::string msg;
                if (!fin_e) {
                    fe = finish_chunk(p, rs, *c);
                    if (fe) msg = last_error();
                }
                std::lock_guard<std::mutex> lk(qm);
                if (fe && !fin_e) {
                    fin_e = fe;
                    fin_msg = msg;
                }
                --pending;
                qcv.notify_all();
            }
            release_workspaces();
        });
        for (auto& chp : runs) {
            ChunkRun& ch = *chp;
            e = solve_chunk(p, rs, ch, ch.group >= 0 ? groups[ch.group].get() : nullptr, shared);
            std::unique_lock<std::mutex> lk(qm);
            if (e || fin_e) break;
            qcv.wait(lk, [&] { return pending == 0 || fin_e != 0; });  // at most one chunk being finished
            if (fin_e) break;
            queue.push_back(chp.get());
            ++pending;
            qcv.notify_all();
        }
        {
            std::lock_guard<std::mutex> lk(qm);
            closed = true;
            qcv.notify_all();
        }
        helper.join();
        set_thread_capture(capture);
        if (e) return e;
        if (fin_e) return set_error("%s", fin_msg.c_str());
    }
    for (size_t i = 0; i < rs.size(); ++i) {
        if (copy_of[i] < 0) continue;
        const ReqState& src = rs[copy_of[i]];
        ReqState& dst = rs[i];
        *dst.result = *src.result;
        const int K = std::max(0, dst.K);
        std::copy(src.beta, src.beta + K, dst.beta);
        std::copy(src.se, src.se + K, dst.se);
        if (dst.vcov) std::copy(src.vcov, src.vcov + (size_t)K * K, dst.vcov);
    }
    return 0;
}

}  // namespace
}  // namespace fehip

using namespace fehip;

extern "C" int fehip_estimate(fehip_panel* panel, const double* d_y, const double* d_W, int K, const double* d_Z,
                              int L, int64_t ld, const uint8_t* d_weights, const fehip_opts* opts,
                              fehip_result* result, double* beta, double* se, double* vcov) {
    (void)hipGetLastError();  // start from a clean error state
    if (!panel) return set_error("fehip_estimate: null argument");
    std::vector<ReqState> rs(1);
    ReqState& s = rs[0];
    s.d_y = d_y;
    s.d_W = d_W;
    s.K = K;
    s.d_Z = d_Z;
    s.L = L;
    s.ld = ld;
    s.d_weights = d_weights;
    s.opts = opts;
    s.result = result;
    s.beta = beta;
    s.se = se;
    s.vcov = vcov;
    const int e = run_batch(reinterpret_cast<Panel*>(panel), rs);
    return e ? e : set_ok();
}

namespace fehip {
int estimate_requests(Panel* p, const fehip_request* reqs, int n_req) {
    std::vector<ReqState> rs(n_req);
    for (int i = 0; i < n_req; ++i) {
        const fehip_request& q = reqs[i];
        ReqState& s = rs[i];
        s.d_y = q.d_y;
        s.d_W = q.d_W;
        s.K = q.K;
        s.d_Z = q.d_Z;
        s.L = q.L;
        s.ld = q.ld;
        s.d_weights = q.d_weights;
        s.opts = &q.opts;
        s.result = q.result;
        s.beta = q.beta;
        s.se = q.se;
        s.vcov = q.vcov;
    }
    return run_batch(p, rs);
}
}  // namespace fehip

extern "C" int fehip_estimate_batch(fehip_panel* panel, const fehip_request* reqs, int n_req) {
    (void)hipGetLastError();
    if (!panel || (n_req > 0 && !reqs) || n_req < 0) return set_error("fehip_estimate_batch: bad arguments");
    const int e = estimate_requests(reinterpret_cast<Panel*>(panel), reqs, n_req);
    return e ? e : set_ok();
}
