// Within-transform in FE-table space (the "CG variant" the reference flags but never implements:
// src/gpu/fe_gpu_demean.f90:63,67 reads use_cg nowhere).
//
// With rows ordered by the primary dim P, the projection onto P's dummies is exact and local to a
// contiguous segment.  Eliminating P leaves the SPD system
//     S a = b,   S = D_o' W M_P D_o,   b = D_o' W M_P x
// over the tables a_o of the other dims o (W = 0/1 row weights).  It is solved by Jacobi-
// preconditioned CG with the exact Schur diagonal, one independent CG per column.  Each CG step
// needs one product S p, computed by streaming only the id arrays:
//     s_i = sum_o p_o[id_o(i)];  u_i = w_i (s_i - mean_{P-segment}(s));  q_o[g] = sum_{i in g} u_i
// Groups of small dims are summed in LDS; larger dims are summed per run of equal ids inside a
// segment and the run partials are reduced in group order (no HBM atomics, fixed order).
// The result M_F x = M_P (x - D_o a) is formed in the final pass over the columns.
//
// Batching: a solve carries up to 32 request slots.  Each row has a 32-bit mask word (bit b set =>
// the row belongs to slot b's subsample) and each column a slot bit, so the columns of many
// requests with the same FE set share every sweep while each keeps its own weights, Schur
// diagonal and CG scalars.  A column's result does not depend on the other columns of its tile.
//
// Stopping rule: max over groups of |r_g / n_g|, the size of the group-mean update the next
// alternating projection would make -- the reference's criterion (fe_gpu_demean.f90:196-211,294).
#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "kstats.hpp"
#include "mfma_gram.hpp"
#include "panel.hpp"
#include "wave.hpp"
#include "within.hpp"

namespace fehip {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxScatter = 6;
constexpr int kMaxLarge = 3;
constexpr int kMaxCT = 16;
constexpr int kMaxCouplingGroups = 64;  // small groups a run-level operator couples (LDS staging)

enum Mode { kIter = 0, kData = 1, kDiag = 2, kFinal = 3 };

struct ScatterDim {
    const int* ids;
    int G;
    int small;
    int64_t tab_off;             // group offset in table space
    int lds_off;                 // small: group offset in LDS
    int large_slot;              // large: 0..kMaxLarge-1
    const int64_t* seg_run_off;  // large
    const int* run_dest;         // large: group-order position of each run
    const int* row_rdest;        // large: per row, run_dest of the run it ends (-1 inside a run)
    double* runbuf;              // large: [n_runs][CT] in group order
    float* runbuf32;             // large: [n_runs][CT] in group order, the FP32 operator's run partials
};

struct ApplyArgs {
    const int64_t* seg_off;
    int64_t GP;
    int elim;
    const uint32_t* wm;        // panel-order row mask words (nullptr: every row belongs to every slot)
    uint32_t cbit[kMaxCT];     // slot bit of each column (0 for padding columns)
    uint32_t anybit;           // OR of the column bits
    uint32_t dbit;             // kDiag: slot bit
    int nd;
    ScatterDim sd[kMaxScatter];
    int Gs;                    // small groups in LDS
    int diag_dim;              // kDiag: the scatter dim processed
    const double* p;           // kIter: [S][CT]
    const int* active;         // kIter: per-column CG activity (nullptr = all active)
    const double* xc[kMaxCT];  // kData/kFinal: input columns (caller row order)
    const float* xc32[kMaxCT]; // kData/kFinal: their FP32 copies, read instead when in32 (wave tiles only)
    int in32;
    double* oc[kMaxCT];        // kFinal: output columns (FP64)
    float* oc32[kMaxCT];       // kFinal: FP32 output columns (panel order), used where non-null
    int out32;                 // kFinal: FP32 output columns for 1 = every column of the tile, 0 = none, 2 = some
    int ncols;                 // valid columns in this tile
    const int* perm;
    double* small_part;        // [blocks][Gs][CT]
    const int64_t* tile_seg;   // [2 * n_tiles] segment range of each block tile
    int64_t n_tiles;
    const int* idP;            // primary ids (panel order): segment heads of the wave tiles
    const int64_t* wtile_seg;  // [2 * n_wtiles] segment range of each wave tile
    const int64_t* wtile_rec;  // [n_wtiles] first row | rows << 32
    int64_t n_wtiles;
    const int64_t* long_segs;  // segments longer than a tile
    int64_t n_long;
    int part_base;             // first partial block written by this launch
    const double* alpha;       // kFinal: solved tables [S][CT]
    int to_work;               // kFinal: 1 = panel-order work columns, 0 = caller columns in place
    double* gram;              // kFinal: per-block 16 x 16 cross-product records (nullptr: none)
    // kIter on a column tile whose columns share one row slot: the lean operator kernel (wave_op_kernel)
    int lean;                  // 1: wave tiles through wave_op_kernel
    uint32_t ubit;             // the tile's slot bit (~0u without row masks)
    int op32;                  // 1: FP32 direction tables (p32) and run partials (runbuf32)
    const float* p32;          // kIter with op32: [S][OpStore<float>::pad(CT)] FP32 copy of p
    double* dump;              // lean kernel: per-wave slots [waves][kMaxCT] that lanes ending no run write
    const RunCoupling* rc;     // kIter: the run-level operator's couplings (host side; nullptr: row kernels)
    const double* dg;          // unweighted solves: Schur diagonal and group counts [S][2]
};

// The final pass holds every row's transformed values of the tile's columns in FP64 registers, so
// the Gram matrix V'V of the tile is accumulated there (mfma_gram.hpp), before any rounding to the
// storage precision.  Each workgroup leaves one record of 256 doubles; records are summed in a fixed
// order.

__device__ __forceinline__ uint32_t row_word(const ApplyArgs& a, int64_t r) { return a.wm ? a.wm[r] : ~0u; }

// v = x (kData/kFinal) and +/- the tables at the row's groups (kIter: +p, kFinal: -alpha)
template <int CT, int MODE>
__device__ __forceinline__ void row_value(const ApplyArgs& a, int64_t r, const double* sp, double (&v)[CT]) {
#pragma unroll
    for (int c = 0; c < CT; ++c) v[c] = 0.0;
    if constexpr (MODE == kData || MODE == kFinal) {
        const int64_t src = a.perm ? a.perm[r] : r;
#pragma unroll
        for (int c = 0; c < CT; ++c) v[c] = c < a.ncols ? a.xc[c][src] : 0.0;
    }
    if constexpr (MODE == kIter || MODE == kFinal) {
        const double sign = MODE == kIter ? 1.0 : -1.0;
        const double* tab = MODE == kIter ? a.p : a.alpha;
        for (int o = 0; o < a.nd; ++o) {
            const ScatterDim& d = a.sd[o];
            const int g = d.ids[r];
            if (d.small) {
                const size_t base = (size_t)(d.lds_off + g) * CT;
#pragma unroll
                for (int c = 0; c < CT; ++c) v[c] += sign * ld_lds(&sp[base + c]);
            } else {
                const int64_t base = (d.tab_off + g) * CT;
#pragma unroll
                for (int c = 0; c < CT; ++c) v[c] += sign * ld_global(&tab[base + c]);
            }
        }
    }
}

// Row data of one wave tile, loaded one tile ahead of its use.
template <int ND>
struct WaveRows {
    int64_t rec;               // first row | rows << 32
    int ip;                    // primary id
    uint32_t word;             // mask word
    int g[ND];                 // ids of the scattered dims
    int rd[ND];                // large dims: run destination of the row (-1 inside a run)
};

// The loads below are issued unconditionally from clamped, always-valid addresses and their values
// selected afterwards: a load under a branch leaves the memory counter with a path-dependent count, and
// the wait for it then drains every load in flight.

// the tile record (first row | rows << 32), 0 past the last tile
__device__ __forceinline__ int64_t load_wave_rec(const ApplyArgs& a, int64_t t) {
    const int64_t v = a.wtile_rec[t < a.n_wtiles ? t : a.n_wtiles - 1];
    return t < a.n_wtiles ? v : 0;
}

// row data of a tile whose record is already in registers, as loaded: lanes past the tile hold the
// first row's values, and the consumer masks them (a select right after a load would wait for it)
template <int ND>
__device__ __forceinline__ void load_wave_rows(const ApplyArgs& a, int64_t rec, int lane, WaveRows<ND>& w) {
    w.rec = rec;
    const int64_t R0 = rec & 0xffffffffll;
    const int nr = (int)(rec >> 32);
    const int64_t r = R0 + (lane < nr ? lane : 0);
    w.ip = a.idP[r];
    w.word = (a.wm ? a.wm : reinterpret_cast<const uint32_t*>(a.idP))[r];
#pragma unroll
    for (int o = 0; o < ND; ++o) {
        // dims past a.nd read the primary ids (valid addresses, values unused)
        w.g[o] = (o < a.nd ? a.sd[o].ids : a.idP)[r];
        w.rd[o] = (o < a.nd ? a.sd[o].row_rdest : a.idP)[r];
    }
}

// A table row of zeros: a row past the tile and a scattered dim past a.nd gather from it, so their
// values need no masking afterwards.
__device__ double g_zero_row[kMaxCT];

// x of lane (addr / 4), with the byte address computed once per tile rather than per column
__device__ __forceinline__ double lane_value(double x, int addr) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)b);
    const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Row values of a wave tile in two phases.  row_loads issues every load of the row without a branch:
// the table entries of all its groups from the global table (small dims too: their few entries stay in
// the L1 cache) and, for kData/kFinal, the input columns, from clamped addresses.  row_combine adds them
// up.  Between the two the caller issues the next tile's prefetch, so the wait for these loads does not
// wait for the prefetch.
template <int CT, int MODE, int ND, typename XT = double>
struct RowLoads {
    double t[(MODE == kIter || MODE == kFinal) ? ND : 1][CT];
    XT x[(MODE == kData || MODE == kFinal) ? CT : 1];
};

// XT: the input columns' type (FP32 copies of registered columns, or the FP64 columns)
template <int CT, int MODE, int ND, typename XT = double>
__device__ __forceinline__ void row_loads(const ApplyArgs& a, int64_t r, const int (&g)[ND],
                                          RowLoads<CT, MODE, ND, XT>& L) {
    if constexpr (MODE == kData || MODE == kFinal) {
        const int64_t src = a.perm ? a.perm[r] : r;
        if constexpr (std::is_same<XT, float>::value) {
#pragma unroll
            for (int c = 0; c < CT; ++c) L.x[c] = ld_global((c < a.ncols ? a.xc32[c] : a.xc32[0]) + src);
        } else {
#pragma unroll
            for (int c = 0; c < CT; ++c) L.x[c] = ld_global((c < a.ncols ? a.xc[c] : a.xc[0]) + src);
        }
    }
    if constexpr (MODE == kIter || MODE == kFinal) {
        const double* tab = MODE == kIter ? a.p : a.alpha;
#pragma unroll
        for (int o = 0; o < ND; ++o) {
            // g < 0: a row past the tile or a dim past a.nd
            const double* src = g[o] >= 0 ? tab + (a.sd[o].tab_off + g[o]) * CT : g_zero_row;
#pragma unroll
            for (int c = 0; c < CT; ++c) L.t[o][c] = ld_global(src + c);
        }
    }
}

template <int CT, int MODE, int ND, typename XT = double>
__device__ __forceinline__ void row_combine(const ApplyArgs& a, const RowLoads<CT, MODE, ND, XT>& L, double (&v)[CT]) {
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        double s = 0.0;
        if constexpr (MODE == kData || MODE == kFinal) s = c < a.ncols ? (double)L.x[c] : 0.0;
        if constexpr (MODE == kIter) {
            s = L.t[0][c];
#pragma unroll
            for (int o = 1; o < ND; ++o) s += L.t[o][c];  // dims past a.nd read the zero row
        }
        if constexpr (MODE == kFinal) {
#pragma unroll
            for (int o = 0; o < ND; ++o) s -= L.t[o][c];
        }
        v[c] = s;
    }
}

// zero the small-dim accumulators and stage the small-dim tables (kIter: p, kFinal: alpha) in LDS
template <int CT, int MODE>
__device__ __forceinline__ void stage_tables(const ApplyArgs& a, double* sacc, double* sp) {
    for (int e = threadIdx.x; e < a.Gs * CT; e += blockDim.x) sacc[e] = 0.0;
    if constexpr (MODE == kIter || MODE == kFinal) {
        const double* tab = MODE == kIter ? a.p : a.alpha;
        for (int o = 0; o < a.nd; ++o) {
            const ScatterDim& d = a.sd[o];
            if (!d.small) continue;
            for (int e = threadIdx.x; e < d.G * CT; e += blockDim.x) sp[(size_t)d.lds_off * CT + e] = tab[d.tab_off * CT + e];
        }
    }
}

template <int CT>
__device__ __forceinline__ bool any_active(const int* active) {
    if (!active) return true;
    bool any = false;
#pragma unroll
    for (int c = 0; c < CT; ++c) any |= active[c] != 0;
    return any;
}

// Run heads of one large dim inside a tile: fills rstart[0..total] (local row of each run start,
// rstart[total] = nr) and returns the number of runs.  Every thread of the block must call it.
__device__ __forceinline__ int tile_runs(bool valid, bool rhead, int j, int nr, int* wcnt, int* rstart) {
    const int lane = j & (kWave - 1), wv = j / kWave;
    const unsigned long long bal = __ballot(rhead);
    if (lane == 0) wcnt[wv] = __popcll(bal);
    __syncthreads();
    int base = 0, total = 0;
    for (int q = 0; q < kBlock / kWave; ++q) {
        base += q < wv ? wcnt[q] : 0;
        total += wcnt[q];
    }
    if (rhead) rstart[base + __popcll(bal & ((1ull << lane) - 1ull))] = j;
    if (j == 0) rstart[total] = nr;
    __syncthreads();
    (void)valid;
    return total;
}

// ------------------------------------------------------------------------------------------------
// Wave tiles: one 64-wide wavefront per tile of whole short primary segments (<= 64 rows in all),
// one row per lane.  Segment sums and run sums are segmented inclusive scans across the lanes in
// registers (DPP row shifts and row broadcasts, a fixed-shape tree), the segment total reaches
// every lane of its segment by one lane permute, and the weighted row counts come from ballots.
// Nothing is staged in LDS except the small-dim tables; the waves of a block run independently.

template <int CT, int MODE, int ND, typename XT = double>
__global__ __launch_bounds__(kBlock) void wave_tile_kernel(ApplyArgs a) {
    // ND: compile-time bound on the scattered dims (a.nd <= ND), so their loops unroll to what is used
    if constexpr (MODE == kIter) {
        if (!any_active<CT>(a.active)) return;
    }
    // slot bit of each column; columns whose CG has stopped are skipped (their table entries are
    // not read again)
    uint32_t cb[CT];
    uint32_t anyb = 0;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        cb[c] = a.cbit[c];
        if constexpr (MODE == kIter)
            if (a.active && !a.active[c]) cb[c] = 0u;
        anyb |= cb[c];
    }
    extern __shared__ double dyn[];
    double* sacc = dyn;                     // [Gs][CT]
    double* sp = dyn + (size_t)a.Gs * CT;   // [Gs][CT]
    stage_tables<CT, MODE>(a, sacc, sp);
    // 1/k for the row counts k = 0..64 of a segment (0 for k = 0), each the correctly rounded quotient,
    // so a lookup gives the value the division would
    __shared__ double inv_count[kWave + 1];
    if (threadIdx.x <= kWave) inv_count[threadIdx.x] = threadIdx.x > 0 ? 1.0 / (double)threadIdx.x : 0.0;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wstride = (int64_t)gridDim.x * (kBlock / kWave);
    gram4_t gacc = {0.0, 0.0, 0.0, 0.0};
    double* gstage = dyn + (size_t)a.Gs * CT * 2;  // kFinal with cross products: per-wave staging
    // software pipeline: the next tile's row data (record, ids, mask words) is loaded while the current
    // tile is processed, so a tile waits only for its table gathers
    int64_t t = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    // records two tiles ahead, row data one tile ahead: loading the next tile's rows never waits for a
    // record requested in the same iteration (gathers a tile ahead as well cost a wave per SIMD of
    // occupancy and more than they hide)
    // the segment scan serves the means (eliminated primary) and the run sums of the large dims; the
    // operator and right-hand-side passes take it unconditionally (their solves nearly always need it,
    // and a branch around it costs a copy of every column at the join)
    const bool need_scan = MODE != kFinal || a.elim != 0;
    WaveRows<ND> cur, nxt;
    load_wave_rows(a, load_wave_rec(a, t), lane, cur);
    int64_t rec_n = load_wave_rec(a, t + wstride);
    for (; t < a.n_wtiles; t += wstride) {
        const int64_t R0 = cur.rec & 0xffffffffll;
        const int nr = (int)(cur.rec >> 32);
        const bool valid = lane < nr;
        const int64_t r = R0 + (valid ? lane : 0);
        const uint32_t word = valid ? (a.wm ? cur.word : ~0u) : 0u;
        const bool on = (word & anyb) != 0;
        const int ip = valid ? cur.ip : -1;
        int gid[ND], rdl[ND];
#pragma unroll
        for (int o = 0; o < ND; ++o) {
            gid[o] = (o < a.nd && valid) ? cur.g[o] : -1;
            rdl[o] = valid ? cur.rd[o] : -1;
        }
        // the tile's loads are requested first; the next tiles' record and rows follow, so waiting for the
        // tile's loads does not wait for the prefetch (one in-order memory counter covers both)
        RowLoads<CT, MODE, ND, XT> ld;
        row_loads<CT, MODE, ND, XT>(a, r, gid, ld);
        const int64_t rec_nn = load_wave_rec(a, t + 2 * wstride);
        load_wave_rows(a, rec_n, lane, nxt);
        rec_n = rec_nn;
        cur = nxt;
        __builtin_amdgcn_sched_barrier(0);  // keep the loads above ahead of their first use
        // every tile holds at least one row (the panel drops tiles of empty segments), so there is no
        // early exit here: a path without this tile's stores would make the wait for the prefetch drain
        double v[CT];
        row_combine<CT, MODE, ND, XT>(a, ld, v);
        // rows outside a column's subsample (and, for the data columns, past the tile) count as zero; an
        // operator product without row masks has nothing to clear (its rows past the tile gathered zeros)
        if (MODE != kIter || a.wm) {
#pragma unroll
            for (int c = 0; c < CT; ++c) v[c] = (on && (word & cb[c])) ? v[c] : 0.0;
        }
        // segment of this lane: [ss, se]
        const int ipp = __shfl_up(ip, 1, kWave);
        const bool hs = valid && (lane == 0 || ip != ipp);
        const unsigned long long H = __ballot(hs);
        const int ss = 63 - __clzll(H & mask_le(lane));
        const unsigned long long nh = H & ~mask_le(lane);
        const int se = nh ? __ffsll((long long)nh) - 2 : nr - 1;
        // One segmented scan per column gives the prefix x = sum of v over [segment start, lane]: its value
        // at the segment end is the segment total (for the mean), and differences of it are the run sums
        // of the large dims, so no second scan is needed.
        double x[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) x[c] = v[c];
        if (need_scan) seg_scan_cols<CT>(x, ScanFactors(ScanMask(lane, ss)));
        const unsigned long long segmask = mask_le(se) & ~((1ull << ss) - 1ull);
        if constexpr (MODE == kFinal) {
            double u[CT];
            if (a.lean) {
                // every column in the slot a.ubit (uniform): one row count, and the segment totals of all
                // columns fetched back to back; without an absorbed primary the mean is 0 (inv = 0)
                const bool ins = (word & a.ubit) != 0;
                const int cnt = __popcll(__ballot(ins) & segmask);
                const double inv = a.elim ? ld_lds(&inv_count[cnt]) : 0.0;
                const double inslot = ins ? 1.0 : 0.0;
                double tot[CT];
#pragma unroll
                for (int c = 0; c < CT; ++c) tot[c] = lane_value(x[c], se << 2);
#pragma unroll
                for (int c = 0; c < CT; ++c) u[c] = fma(-(tot[c] * inv), inslot, v[c]);  // v is 0 outside the slot
            } else {
                double inv = 0.0, inslot = 0.0;
                uint32_t cnt_bit = 0u;
#pragma unroll
                for (int c = 0; c < CT; ++c) {
                    u[c] = 0.0;
                    if (cb[c] == 0u) continue;  // uniform across the wave
                    if (cb[c] != cnt_bit) {     // columns of one slot share the row counts
                        const unsigned long long B = __ballot((word & cb[c]) != 0);
                        const int cnt = __popcll(B & segmask);
                        inv = ld_lds(&inv_count[cnt]);
                        inslot = (word & cb[c]) ? 1.0 : 0.0;
                        cnt_bit = cb[c];
                    }
                    const double mean = a.elim ? lane_value(x[c], se << 2) * inv : 0.0;
                    u[c] = fma(-mean, inslot, v[c]);  // v is 0 outside the slot
                }
            }
            if (valid) {
                // unrolled with a static index into u (an early exit from this loop would index u at run time
                // and serialize the columns' stores behind a scalar load of each column's pointer); the kind of
                // output is uniform over the tile's columns in the common cases, tested once
                if (a.out32 == 1) {
#pragma unroll
                    for (int c = 0; c < CT; ++c)
                        if (c < a.ncols) a.oc32[c][r] = (float)u[c];
                } else if (a.out32 == 0 && a.to_work) {
#pragma unroll
                    for (int c = 0; c < CT; ++c)
                        if (c < a.ncols) a.oc[c][r] = u[c];
                } else {
                    const int64_t src = a.perm ? a.perm[r] : r;
#pragma unroll
                    for (int c = 0; c < CT; ++c) {
                        if (c < a.ncols) {
                            if (a.oc32[c])
                                a.oc32[c][r] = (float)u[c];
                            else if (a.to_work)
                                a.oc[c][r] = u[c];
                            else if (word & cb[c])
                                a.oc[c][src] = u[c];
                        }
                    }
                }
            }
            if (a.gram)  // uniform; u is 0 on rows past the tile and outside each column's slot
                gram_wave<CT>(gstage + (size_t)(threadIdx.x / kWave) * gram_stage<CT>(), u, lane, gacc);
        } else {
            // Per scattered dim: the small dim's LDS row, or the large dim's run geometry.  A new run starts
            // at a segment head or where the id changes; the run sum of u is
            // x(run end) - x(run start - 1) - (rows of the slot in the run) * mean.
            double* tq[ND];
            double* dst[ND];
            int pl_addr[ND];  // byte address of the lane before the run start (its prefix is subtracted)
            const int se_addr = se << 2;
            double keep_pre[ND];
            unsigned long long runmask[ND];
            bool rend[ND];
#pragma unroll
            for (int o = 0; o < ND; ++o) {
                tq[o] = sacc;
                dst[o] = nullptr;
                pl_addr[o] = 0;
                keep_pre[o] = 0.0;
                runmask[o] = 0ull;
                rend[o] = false;
                if (o >= a.nd) continue;
                const ScatterDim& d = a.sd[o];
                const int g = gid[o];
                if (d.small) {
                    tq[o] = sacc + (size_t)(d.lds_off + (on ? g : 0)) * CT;
                    continue;
                }
                const int gp = __shfl_up(g, 1, kWave);
                const bool rh = valid && (hs || g != gp);
                const unsigned long long HR = __ballot(rh);
                const unsigned long long le = HR & mask_le(lane);
                const int rs = 63 - __clzll(le);
                rend[o] = valid && (lane == nr - 1 || (lane < kWave - 1 && ((HR >> (lane + 1)) & 1ull)));
                const int rdest = rdl[o];  // the row's run destination, loaded with its ids
                pl_addr[o] = (rs > 0 ? rs - 1 : 0) << 2;
                runmask[o] = mask_le(lane) & ~((1ull << rs) - 1ull);
                keep_pre[o] = rs == ss ? 0.0 : 1.0;
                dst[o] = d.runbuf + (int64_t)(rend[o] ? rdest : 0) * CT;
            }
            // column by column: mean, the row's value u, its small-dim adds and its run sums, so only the
            // scan prefixes and the row values stay live across the columns
            double inv = 0.0, inslot = 0.0;
            double rcnt[ND];
#pragma unroll
            for (int o = 0; o < ND; ++o) rcnt[o] = 0.0;
            uint32_t cnt_bit = 0u;
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                if (cb[c] == 0u) continue;  // uniform across the wave
                if (cb[c] != cnt_bit) {     // columns of one slot share the row counts
                    const unsigned long long B = __ballot((word & cb[c]) != 0);
                    const int cnt = __popcll(B & segmask);
                    inv = ld_lds(&inv_count[cnt]);
                    inslot = (word & cb[c]) ? 1.0 : 0.0;
                    cnt_bit = cb[c];
#pragma unroll
                    for (int o = 0; o < ND; ++o) rcnt[o] = (double)__popcll(B & runmask[o]);
                }
                const double mean = a.elim ? lane_value(x[c], se_addr) * inv : 0.0;
                const double u = fma(-mean, inslot, v[c]);  // 0 outside the slot and past the tile
#pragma unroll
                for (int o = 0; o < ND; ++o) {
                    if (o >= a.nd) break;
                    if (a.sd[o].small) {
                        // only the lanes in the column's subsample add (a masked-off lane costs no LDS
                        // atomic bandwidth)
                        if (word & cb[c]) lds_add(&tq[o][c], u);
                    } else {
                        const double pre = lane_value(x[c], pl_addr[o]);
                        const double ru = fma(-rcnt[o], mean, fma(-pre, keep_pre[o], x[c]));
                        if (rend[o]) dst[o][c] = ru;
                    }
                }
            }
        }
    }
    if constexpr (MODE != kFinal) {
        __syncthreads();
        double* part = a.small_part + (size_t)(a.part_base + blockIdx.x) * a.Gs * CT;
        for (int e = threadIdx.x; e < a.Gs * CT; e += blockDim.x) part[e] = sacc[e];
    } else {
        if (a.gram) gram_store_block<CT>(gstage, gram_stage<CT>(), gacc, a.gram + (size_t)(a.part_base + blockIdx.x) * 256);
    }
}

// ------------------------------------------------------------------------------------------------
// Lean final pass over the wave tiles, for a column tile whose columns share one row slot and write
// panel-order outputs of one type (the FP32 storage of a mixed-precision solve, or the FP64 work
// columns): u = M_P (x - sum_o alpha_o[g_o]) with the cross products u'u, as the final pass of
// wave_tile_kernel computes them, from a compact argument set.  Column and row addresses are 32-bit
// byte offsets from per-column bases (the host takes this kernel only where every offset fits), so the
// loop keeps its bases in scalar registers and shares one offset register between the columns.
struct FinalArgs {
    const int64_t* wtile_rec;
    int64_t n_wtiles;
    const int* idP;
    const int* perm;            // caller row of each panel row (nullptr: the identity)
    const uint32_t* wm;         // panel-order row mask words (nullptr: every row is in the slot)
    uint32_t ubit;              // the tile's slot bit
    int elim;                   // 1: the primary dim is absorbed (segment means subtracted)
    int ncols;
    const int* ids[3];          // scattered dims: ids (panel order)
    const double* tab[3];       // and their solved tables, alpha + tab_off * CT
    const void* x[kMaxCT];      // input columns (caller row order)
    void* out[kMaxCT];          // output columns (panel order)
    double* gram;               // per-block 16 x 16 cross-product records (nullptr: none)
};

template <int ND>
struct FinalRows {
    int64_t rec;
    int ip, src;
    uint32_t word;
    int g[ND > 0 ? ND : 1];  // ND = 0: only the primary dim is absorbed
};

__device__ __forceinline__ int64_t final_rec(const FinalArgs& a, int64_t t) {
    const int64_t v = a.wtile_rec[t < a.n_wtiles ? t : a.n_wtiles - 1];
    return t < a.n_wtiles ? v : 0;
}

// rows of a tile as loaded, from clamped addresses (lanes past the tile read the first row)
template <int ND>
__device__ __forceinline__ void final_rows(const FinalArgs& a, int64_t rec, int lane, FinalRows<ND>& w) {
    w.rec = rec;
    const uint32_t nr = (uint32_t)(rec >> 32);
    const uint32_t r = (uint32_t)rec + ((uint32_t)lane < nr ? (uint32_t)lane : 0u);
    w.ip = a.idP[r];
    w.src = (a.perm ? a.perm : a.idP)[r];
    w.word = (a.wm ? a.wm : reinterpret_cast<const uint32_t*>(a.idP))[r];
#pragma unroll
    for (int o = 0; o < ND; ++o) w.g[o] = a.ids[o][r];
}

template <typename T>
__device__ __forceinline__ const T* at_offset(const void* base, uint32_t bytes) {
    return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + bytes);
}

template <int CT, int ND, typename XT, typename OT>
__global__ __launch_bounds__(kBlock) void wave_final_kernel(FinalArgs a) {
    extern __shared__ double gst[];  // per-wave staging of the cross products
    __shared__ double inv_count[kWave + 1];
    if (threadIdx.x <= kWave) inv_count[threadIdx.x] = threadIdx.x > 0 ? 1.0 / (double)threadIdx.x : 0.0;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wstride = (int64_t)gridDim.x * (kBlock / kWave);
    gram4_t gacc = {0.0, 0.0, 0.0, 0.0};
    double* mystage = gst + (size_t)(threadIdx.x / kWave) * gram_stage<CT>();
    int64_t t = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    FinalRows<ND> cur, nxt;
    final_rows<ND>(a, final_rec(a, t), lane, cur);
    int64_t rec_n = final_rec(a, t + wstride);
    for (; t < a.n_wtiles; t += wstride) {
        const int nr = (int)(cur.rec >> 32);
        const bool valid = lane < nr;
        const uint32_t r = (uint32_t)cur.rec + (valid ? (uint32_t)lane : 0u);
        const uint32_t src = a.perm ? (uint32_t)cur.src : r;
        const bool ins = valid && (!a.wm || (cur.word & a.ubit) != 0);
        const int ip = valid ? cur.ip : -1;
        // this tile's loads first, then the next tiles' record and rows (the wait for the first does not
        // wait for the prefetch)
        XT xv[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) xv[c] = ld_global(at_offset<XT>(a.x[c < a.ncols ? c : 0], src * (uint32_t)sizeof(XT)));
        double tv[ND > 0 ? ND : 1][CT];
#pragma unroll
        for (int o = 0; o < ND; ++o) {
            const double* row = at_offset<double>(a.tab[o], (uint32_t)cur.g[o] * (uint32_t)(CT * sizeof(double)));
#pragma unroll
            for (int c = 0; c < CT; ++c) tv[o][c] = ld_global(row + c);
        }
        const int64_t rec_nn = final_rec(a, t + 2 * wstride);
        final_rows<ND>(a, rec_n, lane, nxt);
        rec_n = rec_nn;
        cur = nxt;
        __builtin_amdgcn_sched_barrier(0);
        double v[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            double s = (double)xv[c];
#pragma unroll
            for (int o = 0; o < ND; ++o) s -= tv[o][c];
            v[c] = (ins && c < a.ncols) ? s : 0.0;
        }
        // the segment [ss, se] of this lane and its mean (the absorbed primary)
        const int ipp = __shfl_up(ip, 1, kWave);
        const bool hs = valid && (lane == 0 || ip != ipp);
        const unsigned long long H = __ballot(hs);
        const int ss = 63 - __clzll(H & mask_le(lane));
        const unsigned long long nh = H & ~mask_le(lane);
        const int se = nh ? __ffsll((long long)nh) - 2 : nr - 1;
        const unsigned long long segmask = mask_le(se) & ~((1ull << ss) - 1ull);
        double x[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) x[c] = v[c];
        if (a.elim) seg_scan_cols<CT>(x, ScanFactors(ScanMask(lane, ss)));
        const int cnt = __popcll(__ballot(ins) & segmask);
        const double inv = a.elim ? ld_lds(&inv_count[cnt]) : 0.0;
        const double inslot = ins ? 1.0 : 0.0;
        double u[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) u[c] = fma(-(lane_value(x[c], se << 2) * inv), inslot, v[c]);
        if (valid) {
#pragma unroll
            for (int c = 0; c < CT; ++c)
                if (c < a.ncols)
                    *const_cast<OT*>(at_offset<OT>(a.out[c], r * (uint32_t)sizeof(OT))) = (OT)u[c];
        }
        if (a.gram) gram_wave<CT>(mystage, u, lane, gacc);  // u is 0 past the tile and outside the slot
    }
    if (a.gram) gram_store_block<CT>(gst, gram_stage<CT>(), gacc, a.gram + (size_t)blockIdx.x * 256);
}

// Lean operator product q = S p over the wave tiles, for a column tile whose columns share one row slot
// (every solve of a single request).  The slot's row counts, the run geometry of each large dim and the
// segment bounds are formed once per tile rather than once per column, and every column of the tile is
// computed (a column whose CG has stopped costs its share of the work; its output is not read).
// TS is the storage of the direction tables and the run partials:
//   double -- the FP64 operator;
//   float  -- the FP32 operator of the mixed-precision solve: p is gathered from an FP32 copy whose rows
//             are padded to 16 bytes, the row sums, segmented scans (one fused DPP multiply-add per column
//             and step) and run sums are FP32, the small dims' group sums accumulate in FP64 in LDS, and
//             the FP32 run partials are summed per group in FP64 (reduce_runs32_kernel).  Its rounding is
//             corrected by FP64 products of the true residual in the solve around it.
template <typename TS>
struct OpStore {
    // FP32 rows padded to a multiple of 16 bytes (padding them to whole powers of two, so no row straddles
    // a cache line, measured slower: the larger table)
    static constexpr int pad(int ct) { return sizeof(TS) == 4 ? (ct + 3) & ~3 : ct; }
};

__device__ __align__(16) float g_zero_row32[kMaxCT];

__device__ __forceinline__ float lane_value(float x, int addr) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x)));
}

// DATA: the right-hand side b = D_o' W M_P x instead, the rows' values read from the input columns (XT: the
// FP64 columns or the FP32 copies of registered ones; rows in the caller's order, 32-bit byte offsets) and
// every other step as for the operator product, in FP64
template <int CT, int ND, typename TS, bool DATA = false, typename XT = double>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(CT <= 9 && ND <= 2 ? 4 : 1))) void wave_op_kernel(
    ApplyArgs a) {  // four waves per SIMD where that fits without spills (up to 9 columns, 2 scattered dims)
    constexpr bool kF32 = sizeof(TS) == 4;
    static_assert(!DATA || !kF32, "the right-hand side is formed in FP64");
    constexpr int CTP = OpStore<TS>::pad(CT);  // table row stride
    if (!DATA && !any_active<CT>(a.active)) return;
    const TS* tab = kF32 ? reinterpret_cast<const TS*>(a.p32) : reinterpret_cast<const TS*>(a.p);
    const TS* zrow = kF32 ? reinterpret_cast<const TS*>(g_zero_row32) : reinterpret_cast<const TS*>(g_zero_row);
    extern __shared__ double dyn[];
    double* sacc = dyn;  // [Gs][CT] small-dim group sums
    for (int e = threadIdx.x; e < a.Gs * CT; e += blockDim.x) sacc[e] = 0.0;
    __shared__ double inv_count[kWave + 1];
    if (threadIdx.x <= kWave) inv_count[threadIdx.x] = threadIdx.x > 0 ? 1.0 / (double)threadIdx.x : 0.0;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wstride = (int64_t)gridDim.x * (kBlock / kWave);
    const uint32_t ubit = a.ubit;
    int64_t t = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    TS* const wdump = reinterpret_cast<TS*>(a.dump) + t * kMaxCT;
    // the same software pipeline as wave_tile_kernel: records two tiles ahead, rows one tile ahead
    WaveRows<ND> cur, nxt;
    load_wave_rows(a, load_wave_rec(a, t), lane, cur);
    int64_t rec_n = load_wave_rec(a, t + wstride);
    for (; t < a.n_wtiles; t += wstride) {
        const int nr = (int)(cur.rec >> 32);
        const bool valid = lane < nr;
        const uint32_t word = valid ? (a.wm ? cur.word : ~0u) : 0u;
        const bool on = (word & ubit) != 0;
        const int ip = valid ? cur.ip : -1;
        int gid[ND], rdl[ND];
#pragma unroll
        for (int o = 0; o < ND; ++o) {
            gid[o] = (o < a.nd && valid) ? cur.g[o] : -1;
            rdl[o] = cur.rd[o];  // used only on lanes that end a run (valid ones)
        }
        // the tile's gathers (rows past the tile and dims past a.nd read the zero row) or its input columns
        // (rows past the tile read the first row) first, then the next tiles' record and rows
        TS tv[DATA ? 1 : ND][CT];
        XT xv[DATA ? CT : 1];
        if constexpr (DATA) {
            const uint32_t r = (uint32_t)cur.rec + (valid ? (uint32_t)lane : 0u);
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                const void* col;
                if constexpr (sizeof(XT) == 4)
                    col = a.xc32[c < a.ncols ? c : 0];
                else
                    col = a.xc[c < a.ncols ? c : 0];
                xv[c] = ld_global(at_offset<XT>(col, r * (uint32_t)sizeof(XT)));
            }
        } else {
#pragma unroll
            for (int o = 0; o < ND; ++o) {
                const int64_t off = (a.sd[o].tab_off + (gid[o] >= 0 ? gid[o] : 0)) * CTP;
                const TS* src = gid[o] >= 0 ? tab + off : zrow;
#pragma unroll
                for (int c = 0; c < CT; ++c) tv[o][c] = ld_global(src + c);
            }
        }
        const int64_t rec_nn = load_wave_rec(a, t + 2 * wstride);
        load_wave_rows(a, rec_n, lane, nxt);
        rec_n = rec_nn;
        cur = nxt;
        __builtin_amdgcn_sched_barrier(0);
        TS v[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            if constexpr (DATA) {
                v[c] = (on && c < a.ncols) ? (TS)xv[c] : (TS)0;  // rows outside the slot count as zero
            } else {
                TS sum = tv[0][c];
#pragma unroll
                for (int o = 1; o < ND; ++o) sum += tv[o][c];
                v[c] = on ? sum : (TS)0;  // rows outside the slot count as zero
            }
        }
        // segment of this lane: [ss, se]
        const int ipp = __shfl_up(ip, 1, kWave);
        const bool hs = valid && (lane == 0 || ip != ipp);
        const unsigned long long H = __ballot(hs);
        const int ss = 63 - __clzll(H & mask_le(lane));
        const unsigned long long nh = H & ~mask_le(lane);
        const int se = nh ? __ffsll((long long)nh) - 2 : nr - 1;
        TS x[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) x[c] = v[c];
        const ScanMask smask(lane, ss);
        if constexpr (kF32)
            seg_scan_cols_f32<CT>(x, ScanFactorsF(smask));
        else
            seg_scan_cols<CT>(x, ScanFactors(smask));
        const unsigned long long segmask = mask_le(se) & ~((1ull << ss) - 1ull);
        const unsigned long long B = __ballot(on);
        const TS inv = (TS)ld_lds(&inv_count[__popcll(B & segmask)]);
        const int se_addr = se << 2;
        // per scattered dim: the small dim's LDS row, or the large dim's run geometry
        double* tq[ND];
        TS* dst[ND];
        int pl_addr[ND];
        TS keep_pre[ND], rcnt[ND];
        bool rend[ND];
#pragma unroll
        for (int o = 0; o < ND; ++o) {
            tq[o] = sacc;
            dst[o] = nullptr;
            pl_addr[o] = 0;
            keep_pre[o] = rcnt[o] = (TS)0;
            rend[o] = false;
            if (o >= a.nd) continue;
            const ScatterDim& d = a.sd[o];
            const int g = gid[o];
            if (d.small) {
                tq[o] = sacc + (size_t)(d.lds_off + (on ? g : 0)) * CT;
                continue;
            }
            const int gp = __shfl_up(g, 1, kWave);
            const bool rh = valid && (hs || g != gp);
            const unsigned long long HR = __ballot(rh);
            const int rs = 63 - __clzll(HR & mask_le(lane));
            rend[o] = valid && (lane == nr - 1 || (lane < kWave - 1 && ((HR >> (lane + 1)) & 1ull)));
            pl_addr[o] = (rs > 0 ? rs - 1 : 0) << 2;
            const unsigned long long runmask = mask_le(lane) & ~((1ull << rs) - 1ull);
            keep_pre[o] = rs == ss ? (TS)0 : (TS)1;
            rcnt[o] = (TS)__popcll(B & runmask);
            TS* rb = kF32 ? reinterpret_cast<TS*>(d.runbuf32) : reinterpret_cast<TS*>(d.runbuf);
            // lanes ending no run write the wave's dump slot: every lane stores, so every path issues the
            // same memory instructions and the wait for the next tile's prefetched rows never waits for
            // this tile's stores
            dst[o] = rend[o] ? rb + (int64_t)rdl[o] * CT : wdump;
        }
        // segment means of every column (the lane permutes issued back to back), the rows' values u, then
        // per dim the small-dim adds or the run sums; the branches are uniform and outside the column loops
        TS mean[CT];
        if (a.elim) {
#pragma unroll
            for (int c = 0; c < CT; ++c) mean[c] = lane_value(x[c], se_addr);
#pragma unroll
            for (int c = 0; c < CT; ++c) mean[c] *= inv;
        } else {
#pragma unroll
            for (int c = 0; c < CT; ++c) mean[c] = (TS)0;
        }
#pragma unroll
        for (int o = 0; o < ND; ++o) {
            if (o >= a.nd) break;
            if (a.sd[o].small) {
                if (on) {
#pragma unroll
                    for (int c = 0; c < CT; ++c) lds_add(&tq[o][c], (double)(v[c] - mean[c]));
                }
            } else {
                // run sum: x(run end) - x(run start - 1) - (rows of the slot in the run) * mean
                TS pre[CT];
#pragma unroll
                for (int c = 0; c < CT; ++c) pre[c] = lane_value(x[c], pl_addr[o]);
#pragma unroll
                for (int c = 0; c < CT; ++c) dst[o][c] = fma(-rcnt[o], mean[c], fma(-pre[c], keep_pre[o], x[c]));
            }
        }
    }
    __syncthreads();
    double* part = a.small_part + (size_t)(a.part_base + blockIdx.x) * a.Gs * CT;
    for (int e = threadIdx.x; e < a.Gs * CT; e += blockDim.x) part[e] = sacc[e];
}

// ------------------------------------------------------------------------------------------------
// Row tile (kIter, kData, kFinal): one workgroup per tile of whole primary segments (<= kTileRows
// rows), one row per thread.  Rows are staged in LDS; segment means are taken by one thread per
// (segment, column) and run sums by one thread per (run, column), so no thread walks a long serial
// chain and no cross-lane shuffles are needed.  Run partials of a tile are contiguous in the run
// buffer and leave in one coalesced store; small-dim tables take LDS atomics.
template <int CT, int MODE>
__global__ __launch_bounds__(kBlock) void tile_fast_kernel(ApplyArgs a) {
    if constexpr (MODE == kIter) {
        if (!any_active<CT>(a.active)) return;  // every column converged (the host polls periodically)
    }
    __shared__ double sv[kTileRows * CT];   // row values
    __shared__ uint32_t wrow[kTileRows];    // row mask words (0 past the tile end)
    __shared__ int soff[kTileRows + 1];     // local row offsets of the tile's segments
    __shared__ int rstart[kTileRows + 1];   // first row of each run
    __shared__ int wcnt[kBlock / kWave];
    extern __shared__ double dyn[];
    double* sacc = dyn;                     // [Gs][CT]
    double* sp = dyn + (size_t)a.Gs * CT;   // [Gs][CT]
    const int j = threadIdx.x;
    const int lane = j & (kWave - 1);
    gram4_t gacc = {0.0, 0.0, 0.0, 0.0};
    stage_tables<CT, MODE>(a, sacc, sp);
    for (int64_t t = blockIdx.x; t < a.n_tiles; t += gridDim.x) {
        const int64_t s0 = a.tile_seg[2 * t], s1 = a.tile_seg[2 * t + 1];
        const int nseg = (int)(s1 - s0);
        const int64_t R0 = a.seg_off[s0];
        __syncthreads();  // previous tile done with the LDS arrays (and the staged tables are ready)
        for (int k = j; k <= nseg; k += blockDim.x) soff[k] = (int)(a.seg_off[s0 + k] - R0);
        const int nr = (int)(a.seg_off[s1] - R0);
        const bool valid = j < nr;
        const int64_t r = R0 + j;
        const uint32_t word = valid ? row_word(a, r) : 0u;
        const bool on = (word & a.anybit) != 0;
        double v[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) v[c] = 0.0;
        if (on) row_value<CT, MODE>(a, r, sp, v);
#pragma unroll
        for (int c = 0; c < CT; ++c) sv[j * CT + c] = (word & a.cbit[c]) ? v[c] : 0.0;
        wrow[j] = word;
        __syncthreads();
        if (a.elim) {
            // one thread per (segment, column): subtract the column's weighted segment mean
            for (int e = j; e < nseg * CT; e += blockDim.x) {
                const int k = e / CT, c = e - k * CT;
                const uint32_t bit = a.cbit[c];
                const int b = soff[k], f = soff[k + 1];
                double sm = 0.0;
                int nw = 0;
                for (int q = b; q < f; ++q) {
                    sm += sv[q * CT + c];
                    nw += (wrow[q] & bit) ? 1 : 0;
                }
                if (nw > 0) {
                    const double mean = sm / (double)nw;
                    for (int q = b; q < f; ++q)
                        if (wrow[q] & bit) sv[q * CT + c] -= mean;
                }
            }
            __syncthreads();
        }
        if constexpr (MODE == kFinal) {
            if (valid) {
                const int64_t src = a.perm ? a.perm[r] : r;
#pragma unroll
                for (int c = 0; c < CT; ++c) {
                    if (c >= a.ncols) break;
                    if (a.oc32[c])
                        a.oc32[c][r] = (float)sv[j * CT + c];
                    else if (a.to_work)
                        a.oc[c][r] = sv[j * CT + c];
                    else if (word & a.cbit[c])
                        a.oc[c][src] = sv[j * CT + c];
                }
            }
            // rows past the tile end hold 0, and so do values outside each column's slot
            if (a.gram) gram_rows64<CT>(sv + (size_t)(j / kWave) * kWave * CT, lane, gacc);
        } else {
            for (int o = 0; o < a.nd; ++o) {
                const ScatterDim& d = a.sd[o];
                const int g = valid ? d.ids[r] : -1;
                if (d.small) {
                    if (on) {
                        double* tq = sacc + (size_t)(d.lds_off + g) * CT;
#pragma unroll
                        for (int c = 0; c < CT; ++c)
                            if (word & a.cbit[c]) lds_add(&tq[c], sv[j * CT + c]);
                    }
                    continue;
                }
                // runs: a new run starts at a segment start or where the id changes
                int gp = __shfl_up(g, 1, kWave);
                if (lane == 0) gp = (valid && j > 0) ? d.ids[r - 1] : -1;
                bool segstart = false;
                if (valid) {
                    int lo = 0, hi = nseg - 1;  // last segment whose start is <= j
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (soff[mid] <= j)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    segstart = soff[lo] == j;
                }
                const bool rhead = valid && (j == 0 || segstart || g != gp);
                const int total = tile_runs(valid, rhead, j, nr, wcnt, rstart);
                const int64_t rbase = d.seg_run_off[s0];
                for (int e = j; e < total * CT; e += blockDim.x) {
                    const int k = e / CT, c = e - k * CT;
                    double sm = 0.0;
                    for (int q = rstart[k]; q < rstart[k + 1]; ++q) sm += sv[q * CT + c];
                    d.runbuf[(int64_t)d.run_dest[rbase + k] * CT + c] = sm;
                }
                __syncthreads();
            }
        }
    }
    if constexpr (MODE != kFinal) {
        __syncthreads();
        double* part = a.small_part + (size_t)(a.part_base + blockIdx.x) * a.Gs * CT;
        for (int e = j; e < a.Gs * CT; e += blockDim.x) part[e] = sacc[e];
    } else {
        if (a.gram)
            gram_store_block<CT>(dyn + (size_t)a.Gs * CT * 2, gram_stage<CT>(), gacc,
                             a.gram + (size_t)(a.part_base + blockIdx.x) * 256);
    }
}

// ------------------------------------------------------------------------------------------------
// Schur diagonal of one slot and one dim (CT = 2): u0 = w (1 - c/n_seg), where c counts the slot's
// rows of the segment in the same group of diag_dim and n_seg the slot's rows of the segment
// (u0 = w without elimination); u1 = w gives the weighted group counts.
__global__ __launch_bounds__(kBlock) void tile_diag_kernel(ApplyArgs a) {
    constexpr int CT = 2;
    __shared__ double sv[kTileRows * CT];
    __shared__ int gd[kTileRows];
    __shared__ unsigned char wr[kTileRows];
    __shared__ short rs[kTileRows], re[kTileRows];
    __shared__ int soff[kTileRows + 1];
    __shared__ int rstart[kTileRows + 1];
    __shared__ int wcnt[kBlock / kWave];
    extern __shared__ double dyn[];
    double* sacc = dyn;
    const int j = threadIdx.x;
    for (int e = j; e < a.Gs * CT; e += blockDim.x) sacc[e] = 0.0;
    const ScatterDim& d = a.sd[a.diag_dim];
    for (int64_t t = blockIdx.x; t < a.n_tiles; t += gridDim.x) {
        const int64_t s0 = a.tile_seg[2 * t], s1 = a.tile_seg[2 * t + 1];
        const int nseg = (int)(s1 - s0);
        const int64_t R0 = a.seg_off[s0];
        __syncthreads();
        for (int k = j; k <= nseg; k += blockDim.x) soff[k] = (int)(a.seg_off[s0 + k] - R0);
        const int nr = (int)(a.seg_off[s1] - R0);
        const bool valid = j < nr;
        const int64_t r = R0 + j;
        const bool on = valid && (row_word(a, r) & a.dbit);
        gd[j] = valid ? d.ids[r] : -1;
        wr[j] = on ? 1 : 0;
        __syncthreads();
        for (int k = j; k < nseg; k += blockDim.x) {
            const int b = soff[k], e = soff[k + 1];
            for (int q = b; q < e; ++q) {
                rs[q] = (short)b;
                re[q] = (short)e;
            }
        }
        __syncthreads();
        double u0 = 0.0, u1 = 0.0;
        if (on) {
            u1 = 1.0;
            if (a.elim) {
                int cr = 0, nw = 0;
                for (int q = rs[j]; q < re[j]; ++q) {
                    nw += wr[q];
                    cr += (wr[q] && gd[q] == gd[j]) ? 1 : 0;
                }
                u0 = 1.0 - (double)cr / (double)nw;
            } else {
                u0 = 1.0;
            }
        }
        sv[j * CT + 0] = u0;
        sv[j * CT + 1] = u1;
        if (d.small) {
            if (on) {
                lds_add(&sacc[(size_t)(d.lds_off + gd[j]) * CT + 0], u0);
                lds_add(&sacc[(size_t)(d.lds_off + gd[j]) * CT + 1], u1);
            }
            continue;
        }
        const int gp = (valid && j > 0) ? gd[j - 1] : -1;
        const bool rhead = valid && (j == 0 || rs[j] == j || gd[j] != gp);
        const int total = tile_runs(valid, rhead, j, nr, wcnt, rstart);
        const int64_t rbase = d.seg_run_off[s0];
        for (int e = j; e < total * CT; e += blockDim.x) {
            const int k = e / CT, c = e - k * CT;
            double sm = 0.0;
            for (int q = rstart[k]; q < rstart[k + 1]; ++q) sm += sv[q * CT + c];
            d.runbuf[(int64_t)d.run_dest[rbase + k] * CT + c] = sm;
        }
    }
    __syncthreads();
    double* part = a.small_part + (size_t)(a.part_base + blockIdx.x) * a.Gs * CT;
    for (int e = j; e < a.Gs * CT; e += blockDim.x) part[e] = sacc[e];
}

// ------------------------------------------------------------------------------------------------
// Segments longer than a tile: one thread per segment (rare; correctness path).
template <int CT, int MODE>
__global__ __launch_bounds__(kBlock) void long_kernel(ApplyArgs a) {
    if constexpr (MODE == kIter) {
        if (!any_active<CT>(a.active)) return;
    }
    extern __shared__ double dyn[];
    double* sacc = dyn;
    double* sp = dyn + (size_t)a.Gs * CT;
    double* gl = dyn + (size_t)a.Gs * CT * 2;  // kFinal with cross products: the block's 16 x 16 sums
    if constexpr (MODE == kFinal)
        if (a.gram)
            for (int e = threadIdx.x; e < 256; e += blockDim.x) gl[e] = 0.0;
    stage_tables<CT, MODE>(a, sacc, sp);
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; li < a.n_long; li += stride) {
        const int64_t s = a.long_segs[li];
        const int64_t r0 = a.seg_off[s], r1 = a.seg_off[s + 1];
        const uint32_t selbit = MODE == kDiag ? a.dbit : a.anybit;
        double mean[CT];
        int nwc[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            mean[c] = 0.0;
            nwc[c] = 0;
        }
        int nwd = 0;
        if (a.elim) {
            double acc[CT];
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[c] = 0.0;
            for (int64_t r = r0; r < r1; ++r) {
                const uint32_t word = row_word(a, r);
                if (!(word & selbit)) continue;
                ++nwd;
                if constexpr (MODE != kDiag) {
                    double v[CT];
                    row_value<CT, MODE>(a, r, sp, v);
#pragma unroll
                    for (int c = 0; c < CT; ++c)
                        if (word & a.cbit[c]) {
                            acc[c] += v[c];
                            ++nwc[c];
                        }
                }
            }
#pragma unroll
            for (int c = 0; c < CT; ++c) mean[c] = nwc[c] > 0 ? acc[c] / (double)nwc[c] : 0.0;
        }
        if constexpr (MODE == kFinal) {
            for (int64_t r = r0; r < r1; ++r) {
                const uint32_t word = row_word(a, r);
                const int64_t src = a.perm ? a.perm[r] : r;
                double v[CT];
                row_value<CT, MODE>(a, r, sp, v);
                for (int c = 0; c < a.ncols; ++c) {
                    const bool on = word & a.cbit[c];
                    if (a.oc32[c])
                        a.oc32[c][r] = on ? (float)(v[c] - mean[c]) : 0.0f;
                    else if (a.to_work)
                        a.oc[c][r] = on ? v[c] - mean[c] : 0.0;
                    else if (on)
                        a.oc[c][src] = v[c] - mean[c];
                }
                if (a.gram) {  // rare rows: LDS atomics (summation order not fixed)
                    double u[CT];
#pragma unroll
                    for (int c = 0; c < CT; ++c) u[c] = (c < a.ncols && (word & a.cbit[c])) ? v[c] - mean[c] : 0.0;
                    for (int c1 = 0; c1 < a.ncols; ++c1)
                        for (int c2 = 0; c2 < a.ncols; ++c2) {
                            const double pr = u[c1] * u[c2];
                            if (pr != 0.0) lds_add(&gl[c1 * 16 + c2], pr);
                        }
                }
            }
        } else {
            int cur[kMaxLarge];
            int64_t rk[kMaxLarge];
            double racc[kMaxLarge][CT];
#pragma unroll
            for (int k = 0; k < kMaxLarge; ++k) {
                cur[k] = -1;
                rk[k] = 0;
#pragma unroll
                for (int c = 0; c < CT; ++c) racc[k][c] = 0.0;
            }
            for (int o = 0; o < a.nd; ++o)
                if (!a.sd[o].small) rk[a.sd[o].large_slot] = a.sd[o].seg_run_off[s];
            for (int64_t r = r0; r < r1; ++r) {
                const uint32_t word = row_word(a, r);
                double u[CT];
                if constexpr (MODE == kDiag) {
                    const bool on = word & a.dbit;
                    double cr = 0.0;
                    if (on && a.elim) {
                        const int* ids = a.sd[a.diag_dim].ids;
                        const int g = ids[r];
                        for (int64_t q = r0; q < r1; ++q) cr += (ids[q] == g && (row_word(a, q) & a.dbit)) ? 1.0 : 0.0;
                    }
#pragma unroll
                    for (int c = 0; c < CT; ++c) u[c] = 0.0;
                    if (on) {
                        u[0] = a.elim ? 1.0 - cr / (double)nwd : 1.0;
                        if (CT > 1) u[1 % CT] = 1.0;
                    }
                } else {
                    double v[CT];
                    row_value<CT, MODE>(a, r, sp, v);
#pragma unroll
                    for (int c = 0; c < CT; ++c) u[c] = (word & a.cbit[c]) ? v[c] - mean[c] : 0.0;
                }
                for (int o = 0; o < a.nd; ++o) {
                    if (MODE == kDiag && o != a.diag_dim) continue;
                    const ScatterDim& d = a.sd[o];
                    const int g = d.ids[r];
                    if (d.small) {
                        double* t = sacc + (size_t)(d.lds_off + g) * CT;
#pragma unroll
                        for (int c = 0; c < CT; ++c)
                            if (u[c] != 0.0) lds_add(&t[c], u[c]);
                    } else {
#pragma unroll
                        for (int k = 0; k < kMaxLarge; ++k) {
                            if (k != d.large_slot) continue;
                            if (g != cur[k]) {
                                if (cur[k] >= 0) {
                                    double* dst = d.runbuf + (int64_t)d.run_dest[rk[k]] * CT;
#pragma unroll
                                    for (int c = 0; c < CT; ++c) dst[c] = racc[k][c];
                                    ++rk[k];
                                }
                                cur[k] = g;
#pragma unroll
                                for (int c = 0; c < CT; ++c) racc[k][c] = 0.0;
                            }
#pragma unroll
                            for (int c = 0; c < CT; ++c) racc[k][c] += u[c];
                        }
                    }
                }
            }
            for (int o = 0; o < a.nd; ++o) {
                if (MODE == kDiag && o != a.diag_dim) continue;
                const ScatterDim& d = a.sd[o];
                if (d.small) continue;
#pragma unroll
                for (int k = 0; k < kMaxLarge; ++k) {
                    if (k != d.large_slot || cur[k] < 0) continue;
                    double* dst = d.runbuf + (int64_t)d.run_dest[rk[k]] * CT;
#pragma unroll
                    for (int c = 0; c < CT; ++c) dst[c] = racc[k][c];
                }
            }
        }
    }
    if constexpr (MODE != kFinal) {
        __syncthreads();
        double* part = a.small_part + (size_t)(a.part_base + blockIdx.x) * a.Gs * CT;
        for (int e = threadIdx.x; e < a.Gs * CT; e += blockDim.x) part[e] = sacc[e];
    } else {
        if (a.gram) {
            __syncthreads();
            double* rec = a.gram + (size_t)(a.part_base + blockIdx.x) * 256;
            for (int e = threadIdx.x; e < 256; e += blockDim.x) rec[e] = gl[e];
        }
    }
}

// Sum of the per-block cross-product records in a fixed order: stage 1 adds runs of kGramRun records,
// stage 2 (one block) adds the stage-1 sums and writes the nc x nc corner into out (leading dim ld).
constexpr int kGramRun = 32;

__global__ void gram_sum_kernel(const double* __restrict__ rec, int nrec, double* __restrict__ part) {
    const int e = threadIdx.x;  // 256 threads, one record element each
    const int r0 = blockIdx.x * kGramRun, r1 = min(nrec, r0 + kGramRun);
    double t = 0.0;
    for (int r = r0; r < r1; ++r) t += rec[(size_t)r * 256 + e];
    part[(size_t)blockIdx.x * 256 + e] = t;
}

__global__ void gram_final_kernel(const double* __restrict__ part, int nblk, int nc, double* out, int64_t ld) {
    const int e = threadIdx.x;
    double t = 0.0;
    for (int b = 0; b < nblk; ++b) t += part[(size_t)b * 256 + e];
    const int i = e >> 4, j = e & 15;
    if (i < nc && j < nc) out[(int64_t)j * ld + i] = t;
}

// ------------------------------------------------------------------------------------------------
// q[g][c] = sum of the run partials of group g (runs in group order)
template <int CT>
__global__ void reduce_runs_kernel(const double* __restrict__ runbuf, int sp, const int64_t* __restrict__ goff, int G,
                                   double* q, const int* __restrict__ active) {
    if (!any_active<CT>(active)) return;
    const int64_t total = (int64_t)G * CT;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int g = (int)(e / CT), c = (int)(e % CT);
        // four independent partial sums keep four loads in flight per thread
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        const int64_t k0 = goff[g], k1 = goff[g + 1];
        int64_t k = k0;
        for (; k + 4 <= k1; k += 4) {
            s0 += runbuf[k * sp + c];
            s1 += runbuf[(k + 1) * sp + c];
            s2 += runbuf[(k + 2) * sp + c];
            s3 += runbuf[(k + 3) * sp + c];
        }
        for (; k < k1; ++k) s0 += runbuf[k * sp + c];
        q[e] = (s0 + s1) + (s2 + s3);
    }
}

// q[g][c] = FP64 sum of the FP32 run partials of group g (runs in group order)
template <int CT>
__global__ void reduce_runs32_kernel(const float* __restrict__ runbuf, int sp, const int64_t* __restrict__ goff, int G,
                                     double* q, const int* __restrict__ active) {
    if (!any_active<CT>(active)) return;
    const int64_t total = (int64_t)G * CT;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int g = (int)(e / CT), c = (int)(e % CT);
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        const int64_t k0 = goff[g], k1 = goff[g + 1];
        int64_t k = k0;
        for (; k + 4 <= k1; k += 4) {
            s0 += (double)runbuf[k * sp + c];
            s1 += (double)runbuf[(k + 1) * sp + c];
            s2 += (double)runbuf[(k + 2) * sp + c];
            s3 += (double)runbuf[(k + 3) * sp + c];
        }
        for (; k < k1; ++k) s0 += (double)runbuf[k * sp + c];
        q[e] = (s0 + s1) + (s2 + s3);
    }
}

// one workgroup per table entry: the threads stride over the per-block partials in a fixed order and
// combine them with a fixed-shape tree
template <int CT>
__global__ __launch_bounds__(kBlock) void reduce_small_kernel(const double* __restrict__ part, int nblk, int Gs,
                                                               int lds_off, int G, double* q) {
    __shared__ double ws[kBlock / kWave];
    const int e = blockIdx.x;
    if (e >= G * CT) return;
    const size_t off = (size_t)lds_off * CT + e;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(size_t)b * Gs * CT + off];
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += ws[w];
        q[e] = t;
    }
}

// ---- run-level operator ------------------------------------------------------------------------
// An unweighted solve that absorbs the primary dim P, one large dim D and small dims s (Gs groups in
// all) has S p = D_o' M_P D_o p = (A p_D + B p_s, B' p_D + C p_s) with A = D_D' M_P D_D and the
// couplings B = D_D' M_P D_s, C = D_s' M_P D_s built once (RunCoupling).  A needs no row: in a primary
// segment of n rows whose runs of D (group g_r, length l_r) give m = sum_r l_r p_D[g_r] / n, run r
// contributes l_r (p_D[g_r] - m) to group g_r.

// A run partial of the run-level operator.  FP64: CT values padded to a 16-byte multiple and written with
// 16-byte stores (the partials land at scattered places, so every store instruction touches a line per
// lane: fewer, wider stores touch fewer; c2: operator 1.36 -> 1.30 ms).  FP32: CT values unpadded (the
// padded 16-byte form measured slower there, 0.97 -> 1.05 ms, its larger rows outweighing the wider stores).
template <int CT, typename TS>
struct RunRow {
    static constexpr int kStride = sizeof(TS) == 4 ? CT : (CT + 1) & ~1;
    __device__ __forceinline__ static void store(TS* dst, const TS (&w)[kStride]) {
        if constexpr (sizeof(TS) == 4) {
#pragma unroll
            for (int c = 0; c < kStride; ++c) dst[c] = w[c];
        } else {
#pragma unroll
            for (int c = 0; c < kStride; c += 2) *reinterpret_cast<double2*>(dst + c) = double2{w[c], w[c + 1]};
        }
    }
};

// One wave per run tile (whole primary segments, at most 64 runs): a lane per run gathers its group's
// row of p_D, and a segmented scan of l p_D and of l over the tile's segments gives each segment's mean m,
// which the segment's last lane writes at the segment's place ([G_P] rows, one contiguous piece per wave).
// The group sums A p_D = sum_r l_r (p_D[g_r] - m_seg(r)) are then formed per group from the group-order list
// of runs (reduce_segmeans_kernel), gathering m by segment: 5M segment rows at c2 instead of 19.5M run
// partials (c2 operator product 0.76 + 0.68 ms with the run partials written in run order and gathered per
// group; 1.30 + 0.36 ms with them written at their group-order places).  Table rows have stride ctp; bit 31
// of run_len marks a segment's first run.
template <int CT, typename TS>
__global__ __launch_bounds__(kBlock) void run_wave_kernel(const TS* __restrict__ tab, int ctp, int64_t tab_off,
                                                          const int64_t* __restrict__ rtile_rec,
                                                          const int64_t* __restrict__ rtile_seg, int64_t n_rtiles,
                                                          const int* __restrict__ run_gid,
                                                          const int* __restrict__ run_len, TS* __restrict__ mseg,
                                                          const int* __restrict__ active) {
    if (!any_active<CT>(active)) return;
    constexpr bool kF32 = sizeof(TS) == 4;
    constexpr int SP = RunRow<CT, TS>::kStride;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wstride = (int64_t)gridDim.x * (kBlock / kWave);
    for (int64_t t = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; t < n_rtiles; t += wstride) {
        const int64_t rec = rtile_rec[t];
        const int64_t seg0 = rtile_seg[t];
        const int64_t r0 = (int64_t)(uint32_t)rec;
        const int nr = (int)(rec >> 32);
        const bool valid = lane < nr;
        const int64_t r = r0 + (valid ? lane : 0);
        const int lw = run_len[r];
        const int g = run_gid[r];
        const bool head = valid && (lane == 0 || lw < 0);
        const TS l = valid ? (TS)(lw & 0x7fffffff) : (TS)0;
        const TS* row = tab + (tab_off + g) * ctp;
        TS x[CT + 1];
#pragma unroll
        for (int c = 0; c < CT; ++c) x[c] = l * ld_global(row + c);
        x[CT] = l;
        const unsigned long long H = __ballot(head);
        const unsigned long long nh = H & ~mask_le(lane);
        const int ss = 63 - __clzll(H & mask_le(lane));
        const int se = nh ? __ffsll((long long)nh) - 2 : nr - 1;
        const ScanMask smask(lane, ss);
        if constexpr (kF32)
            seg_scan_cols_f32<CT + 1>(x, ScanFactorsF(smask));
        else
            seg_scan_cols<CT + 1>(x, ScanFactors(smask));
        if (valid && lane == se) {  // the segment's last run holds its totals (no rows in a subsample: m = 0)
            const TS inv = x[CT] > (TS)0 ? (TS)1 / x[CT] : (TS)0;
            TS w[SP];
#pragma unroll
            for (int c = 0; c < SP; ++c) w[c] = c < CT ? x[c] * inv : (TS)0;
            const int64_t seg = seg0 + __popcll(H & mask_le(lane)) - 1;
            RunRow<CT, TS>::store(mseg + seg * SP, w);
        }
    }
}

// Segments of more than 64 runs: one thread per listed segment writes its mean.
template <int CT, typename TS>
__global__ __launch_bounds__(kBlock) void run_long_kernel(const TS* __restrict__ tab, int ctp, int64_t tab_off,
                                                          const int64_t* __restrict__ segs, int64_t nseg,
                                                          const int64_t* __restrict__ seg_run_off,
                                                          const int* __restrict__ run_gid,
                                                          const int* __restrict__ run_len, TS* __restrict__ mseg,
                                                          const int* __restrict__ active) {
    if (!any_active<CT>(active)) return;
    constexpr int SP = RunRow<CT, TS>::kStride;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nseg; i += stride) {
        const int64_t w = segs[i];
        const int64_t r0 = seg_run_off[w], r1 = seg_run_off[w + 1];
        double m[CT], nw = 0.0;
#pragma unroll
        for (int c = 0; c < CT; ++c) m[c] = 0.0;
        for (int64_t r = r0; r < r1; ++r) {
            const TS* row = tab + (tab_off + run_gid[r]) * ctp;
            const double l = (double)(run_len[r] & 0x7fffffff);
            nw += l;
#pragma unroll
            for (int c = 0; c < CT; ++c) m[c] += l * (double)row[c];
        }
        const double inv = nw > 0.0 ? 1.0 / nw : 0.0;
        TS v[SP];
#pragma unroll
        for (int c = 0; c < SP; ++c) v[c] = c < CT ? (TS)(m[c] * inv) : (TS)0;
        RunRow<CT, TS>::store(mseg + w * SP, v);
    }
}

// q[g][c] = sum over the runs r of group g (group order) of l_r (p_D[g][c] - m[seg(r)][c]): the run-level
// operator's A p_D from the segment means, gathered by segment; FP64 arithmetic whatever the tables' type
template <int CT, typename TS>
__global__ void reduce_segmeans_kernel(const TS* __restrict__ tab, int ctp, int64_t tab_off,
                                       const TS* __restrict__ mseg, const int* __restrict__ grp_seg,
                                       const int* __restrict__ grp_len, const int64_t* __restrict__ goff, int G,
                                       double* q, const int* __restrict__ active) {
    if (!any_active<CT>(active)) return;
    constexpr int SP = RunRow<CT, TS>::kStride;
    const int64_t total = (int64_t)G * CT;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int g = (int)(e / CT), c = (int)(e % CT);
        const double pg = (double)tab[(tab_off + g) * ctp + c];
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        const int64_t k0 = goff[g], k1 = goff[g + 1];
        int64_t k = k0;
        for (; k + 4 <= k1; k += 4) {
            const int64_t a0 = grp_seg[k], a1 = grp_seg[k + 1], a2 = grp_seg[k + 2], a3 = grp_seg[k + 3];
            const double l0 = grp_len[k], l1 = grp_len[k + 1], l2 = grp_len[k + 2], l3 = grp_len[k + 3];
            s0 += l0 * (pg - (double)mseg[a0 * SP + c]);
            s1 += l1 * (pg - (double)mseg[a1 * SP + c]);
            s2 += l2 * (pg - (double)mseg[a2 * SP + c]);
            s3 += l3 * (pg - (double)mseg[a3 * SP + c]);
        }
        for (; k < k1; ++k) s0 += (double)grp_len[k] * (pg - (double)mseg[(int64_t)grp_seg[k] * SP + c]);
        q[e] = (s0 + s1) + (s2 + s3);
    }
}

// segment of every run (runs in panel order)
__global__ void run_segments_kernel(const int64_t* __restrict__ seg_run_off, int64_t GP, int* __restrict__ run_seg) {
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < GP; w += (int64_t)gridDim.x * blockDim.x)
        for (int64_t r = seg_run_off[w]; r < seg_run_off[w + 1]; ++r) run_seg[r] = (int)w;
}

// the runs in group order: segment and length of each
__global__ void group_order_runs_kernel(const int* __restrict__ run_order, const int* __restrict__ run_seg,
                                        const int* __restrict__ run_len, int64_t R, int* __restrict__ grp_seg,
                                        int* __restrict__ grp_len) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < R; k += (int64_t)gridDim.x * blockDim.x) {
        const int r = run_order[k];
        grp_seg[k] = run_seg[r];
        grp_len[k] = run_len[r] & 0x7fffffff;
    }
}

constexpr int kCouplingTile = 64;  // groups of D per LDS tile of the coupling pass
typedef double double4_t __attribute__((ext_vector_type(4)));

// One pass over B in tiles of kCouplingTile groups of D (B's rows and p_D's rows staged in LDS, zero past
// Gs and CT) on the FP64 matrix cores (v_mfma_f64_16x16x4f64: A element (i, k) in lane i + 16 k, B
// element (k, j) in lane j + 16 k, D[i][j] in lane j + 16 (i % 4), element i / 4): wave w adds
// (B p_s) of the tile's rows 16 w .. 16 w + 15 to q_D, and accumulates its quarter of the tile's
// reduction in B' p_D, kept across tiles; the block's four quarters are summed at the end into
// part[blk][t][c].
template <int CT, typename TS>
__global__ __launch_bounds__(kBlock) void coupling_pass_kernel(const TS* __restrict__ B, int Gs, int G,
                                                               int64_t tab_off, const TS* __restrict__ tab, int ctp,
                                                               const int* __restrict__ small_rows, double* q,
                                                               double* part, const int* __restrict__ active) {
    static_assert(kCouplingTile == 4 * 16 && kMaxCouplingGroups == 64 && kMaxCT <= 16, "coupling tile shape");
    if (!any_active<CT>(active)) return;
    __shared__ double sb[kCouplingTile * kMaxCouplingGroups];  // the tile's rows of B, row stride Gs
    __shared__ double sp[kCouplingTile][16 + 1];                // [g][c]
    __shared__ double ps[kMaxCouplingGroups][16 + 1];           // p_s: [t][c], zero past Gs and CT
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int lr = lane & 15, lk = lane >> 4;
    const int KS1 = (Gs + 3) / 4;   // k-steps over t (B p_s)
    const int MB = (Gs + 15) / 16;  // 16-row blocks of t (B' p_D)
    for (int e = threadIdx.x; e < kMaxCouplingGroups * 16; e += blockDim.x) {
        const int t = e / 16, c = e % 16;
        ps[t][c] = (t < Gs && c < CT) ? (double)tab[(int64_t)small_rows[t] * ctp + c] : 0.0;
    }
    double4_t acc2[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) acc2[mb] = (double4_t){0.0, 0.0, 0.0, 0.0};
    constexpr int kBPer = kCouplingTile * kMaxCouplingGroups / kBlock, kPPer = kCouplingTile * 16 / kBlock;
    for (int64_t g0 = (int64_t)blockIdx.x * kCouplingTile; g0 < G; g0 += (int64_t)gridDim.x * kCouplingTile) {
        const int ng = (int)(G - g0 < kCouplingTile ? G - g0 : kCouplingTile);
        const int nb = ng * Gs;  // the tile's entries of B, contiguous
        // every load of the tile is issued (clamped addresses, no branch) before the LDS stores wait on them
        double rb[kBPer], rp[kPPer];
#pragma unroll
        for (int k = 0; k < kBPer; ++k) {
            const int e = threadIdx.x + kBlock * k;
            const double v = (double)B[g0 * Gs + (e < nb ? e : 0)];
            rb[k] = e < nb ? v : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kPPer; ++k) {
            const int e = threadIdx.x + kBlock * k;
            const int gi = e / 16, c = e % 16;
            const bool ok = gi < ng && c < CT;
            const double v = (double)tab[ok ? (tab_off + g0 + gi) * ctp + c : 0];
            rp[k] = ok ? v : 0.0;
        }
        __syncthreads();  // the previous tile's reads are done
#pragma unroll
        for (int k = 0; k < kBPer; ++k) sb[threadIdx.x + kBlock * k] = rb[k];
#pragma unroll
        for (int k = 0; k < kPPer; ++k) {
            const int e = threadIdx.x + kBlock * k;
            sp[e / 16][e % 16] = rp[k];
        }
        __syncthreads();
        // q_D rows 16 w + i: (B tile)(p_s); B entries past Gs or past the tile's rows read as zero
        double4_t acc = {0.0, 0.0, 0.0, 0.0};
        const int gr = w * 16 + lr;
#pragma unroll
        for (int ks = 0; ks < 16; ++ks)
            if (ks < KS1) {
                const int t = ks * 4 + lk;
                const double a = (gr < ng && t < Gs) ? sb[gr * Gs + t] : 0.0;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, ps[t][lr], acc, 0, 0, 0);
            }
        if (lr < CT) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = w * 16 + lk + 4 * v;
                if (gi < ng) q[(tab_off + g0 + gi) * CT + lr] += acc[v];
            }
        }
        // (B tile)' (p_D tile) over this wave's 16 groups of the tile
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
            if (mb < MB) {
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    const int gi = w * 16 + ks * 4 + lk, t = mb * 16 + lr;
                    const double a = (gi < ng && t < Gs) ? sb[gi * Gs + t] : 0.0;
                    acc2[mb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sp[gi][lr], acc2[mb], 0, 0, 0);
                }
            }
    }
    // the four waves' quarters, summed in wave order
    __syncthreads();
    double* red = sb;  // [4 waves][4 mb][64 lanes][4]: 4096 doubles
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[((w * 4 + mb) * kWave + lane) * 4 + v] = acc2[mb][v];
    __syncthreads();
    for (int pi = threadIdx.x; pi < Gs * CT; pi += blockDim.x) {
        const int t = pi / CT, c = pi % CT;
        const int mb = t / 16, i = t % 16;
        const int ln = c + 16 * (i % 4), v = i / 4;
        double s2 = 0.0;
        for (int ww = 0; ww < kBlock / kWave; ++ww) s2 += red[((ww * 4 + mb) * kWave + ln) * 4 + v];
        part[(int64_t)blockIdx.x * Gs * CT + pi] = s2;
    }
}

// q_s[t][c] = the blocks' parts summed in a fixed order + sum_t2 C[t][t2] p_s[t2][c]: one workgroup per
// entry (t, c), written to the small group's table row
template <int CT, typename TS>
__global__ __launch_bounds__(kBlock) void coupling_final_kernel(const double* __restrict__ part, int nblk,
                                                                const double* __restrict__ C, int Gs,
                                                                const TS* __restrict__ tab, int ctp,
                                                                const int* __restrict__ small_rows, double* q,
                                                                const int* __restrict__ active) {
    if (!any_active<CT>(active)) return;
    __shared__ double ws[kBlock / kWave];
    const int pi = blockIdx.x;
    const int t = pi / CT, c = pi % CT;
    double s2 = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s2 += part[(int64_t)b * Gs * CT + pi];
    for (int t2 = threadIdx.x; t2 < Gs; t2 += blockDim.x)
        s2 += C[t * Gs + t2] * (double)tab[(int64_t)small_rows[t2] * ctp + c];
    s2 = wave_sum(s2);
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tt = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) tt += ws[w];
        q[(int64_t)small_rows[t] * CT + c] = tt;
    }
}

// Run lengths over one row slot's subsample: the row that ends each run (row_rdest >= 0) counts the run's rows
// in the slot, walking back over the run (rows of equal primary and D ids); written in group order (grp_lw)
// and at the run's place with the segment-head bit of run_len (lw).
__global__ void weighted_runs_kernel(const int* __restrict__ idP, const int* __restrict__ idD,
                                     const int* __restrict__ row_rdest, const uint32_t* __restrict__ wm, uint32_t bit,
                                     const int* __restrict__ run_order, const int* __restrict__ run_len, int64_t n,
                                     int* __restrict__ grp_lw, int* __restrict__ lw) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int k = row_rdest[i];
        if (k < 0) continue;
        const int p0 = idP[i], d0 = idD[i];
        int c = 0;
        for (int64_t j = i; j >= 0 && idP[j] == p0 && idD[j] == d0; --j) c += (wm[j] & bit) ? 1 : 0;
        grp_lw[k] = c;
        const int r = run_order[k];
        lw[r] = c | (run_len[r] & (int)0x80000000);
    }
}

// The same counts over the wave tiles (every primary segment in one): a wave per tile, a lane per row; the rows
// of the slot in a run are a popcount of the slot's ballot over the run's lanes, formed at the run's last lane.
__global__ __launch_bounds__(kBlock) void weighted_runs_wave_kernel(const int64_t* __restrict__ wtile_rec,
                                                                    int64_t n_wtiles, const int* __restrict__ row_rdest,
                                                                    const uint32_t* __restrict__ wm, uint32_t bit,
                                                                    const int* __restrict__ run_order,
                                                                    const int* __restrict__ run_len,
                                                                    int* __restrict__ grp_lw, int* __restrict__ lw) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wstride = (int64_t)gridDim.x * (kBlock / kWave);
    for (int64_t t = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; t < n_wtiles; t += wstride) {
        const int64_t rec = wtile_rec[t];
        const int nr = (int)(rec >> 32);
        const bool valid = lane < nr;
        const int64_t r = (rec & 0xffffffffll) + (valid ? lane : 0);
        const int k = valid ? row_rdest[r] : -1;  // >= 0 at the last row of a run
        const bool in = valid && (wm[r] & bit) != 0;
        const unsigned long long E = __ballot(k >= 0);  // run ends
        const unsigned long long B = __ballot(in);
        // the run of this lane starts after the previous run end
        const unsigned long long before = E & ((1ull << lane) - 1ull);
        const int rs = before ? 64 - __clzll(before) : 0;
        if (k >= 0) {
            const int c = __popcll(B & mask_le(lane) & ~((1ull << rs) - 1ull));
            grp_lw[k] = c;
            const int ro = run_order[k];
            lw[ro] = c | (run_len[ro] & (int)0x80000000);
        }
    }
}

// out = (float) in, n entries
__global__ void to_f32_kernel(const double* __restrict__ in, int64_t n, float* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

// unit directions of small groups t0 .. t0 + nj - 1: P[small_rows[t0 + j]][j] = 1 (P cleared before)
__global__ void coupling_units_kernel(double* P, int ct, const int* __restrict__ small_rows, int t0, int nj) {
    const int j = threadIdx.x;
    if (j < nj) P[(int64_t)small_rows[t0 + j] * ct + j] = 1.0;
}

// B[g][t0 + j] and C[t][t0 + j] from the operator's products Q ([S][ct]) of those unit directions
__global__ void coupling_extract_kernel(const double* __restrict__ Q, int ct, int t0, int nj, int64_t tab_off, int G,
                                        int Gs, const int* __restrict__ small_rows, double* B, double* C) {
    const int64_t total = (int64_t)(G + Gs) * nj;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int64_t i = e / nj;
        const int j = (int)(e - i * nj);
        if (i < G)
            B[i * Gs + t0 + j] = Q[(tab_off + i) * ct + j];
        else
            C[(i - G) * Gs + t0 + j] = Q[(int64_t)small_rows[i - G] * ct + j];
    }
}

// q_D[g][c] = n_g p_D[g][c]: with the primary dim not absorbed every row lies in one group of D, so
// A = D_D' D_D is the diagonal of D's group counts (the first entry of the Schur diagonal's rows)
template <int CT>
__global__ __launch_bounds__(kBlock) void diag_scale_kernel(const double* __restrict__ dg,
                                                            const double* __restrict__ p, int64_t tab_off, int G,
                                                            double* __restrict__ q, const int* __restrict__ active) {
    if (!any_active<CT>(active)) return;
    const int64_t total = (int64_t)G * CT;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int64_t g = e / CT;
        q[tab_off * CT + e] = dg[(tab_off + g) * 2] * p[tab_off * CT + e];
    }
}

// ---- CG in table space ([S][CT], one CG per column) ---------------------------------------------
// tol_in: the level at which a round of a column's CG is done -- tol, or in a round of the mixed-precision
// solve max(tol, delta * the criterion of the round's true residual).  A column stops (active = 0) once
// its criterion is below tol; below tol_in only `reached` is set and the column goes on iterating until
// every column of the tile has reached its level, so no column waits idle for the slowest.
struct CgState {
    double rz[kMaxCT], alpha[kMaxCT], beta[kMaxCT], maxn[kMaxCT], tol_in[kMaxCT];
    int active[kMaxCT], iters[kMaxCT], reached[kMaxCT];
};

// Flat element loops over a [S][CT] table for the CG vector kernels: thread t of the grid takes elements
// t, t + T, t + 2T, ... with T the largest multiple of CT not above the grid's thread count, so each thread
// stays on one column (t % CT) and a wave's loads are consecutive elements (one 512-byte piece per FP64 load
// instruction, where a thread per table row reads rows CT elements apart).
template <int CT>
struct FlatCols {
    int64_t t, T, n;
    int c;
    __device__ __forceinline__ explicit FlatCols(int64_t S) {
        const int64_t all = (int64_t)gridDim.x * blockDim.x;
        T = all - all % CT;
        t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        c = (int)(t % CT);
        n = S * CT;
    }
    __device__ __forceinline__ bool on() const { return t < T; }
};

// Column partials of a block (sum of a, max of m) at part_a[block][c] and part_m[block][c]: thread i of the
// block holds column (c0 + i) % CT, with c0 the column of its first thread; threads past T hold zeros.  Every
// thread of the block calls it; the order of the sums is fixed.
template <int CT>
__device__ __forceinline__ void col_partials(double a, double m, double* part_a, double* part_m) {
    __shared__ double sa[kBlock], sm[kBlock];
    sa[threadIdx.x] = a;
    sm[threadIdx.x] = m;
    __syncthreads();
    if (threadIdx.x < CT) {
        const int c0 = (int)(((int64_t)blockIdx.x * blockDim.x) % CT);
        const int c = threadIdx.x;
        double ta = 0.0, tm = 0.0;
        for (int i = (c - c0 + CT) % CT; i < (int)blockDim.x; i += CT) {
            ta += sa[i];
            tm = fmax(tm, sm[i]);
        }
        part_a[blockIdx.x * CT + c] = ta;
        if (part_m) part_m[blockIdx.x * CT + c] = tm;
    }
}

// CG start: z = M^-1 r, p = z (and its FP32 copy), x = 0.  With b and q given (the reliable update of
// the mixed-precision solve), r = b - q, the true residual of the current x, and z = M^-1 r only: x and
// the direction p are kept (cg_reliable_scalars and cg_pdir_kernel continue the same CG).
template <int CT>
__global__ __launch_bounds__(kBlock) void cg_init_kernel(int64_t S, double* __restrict__ r,
                                                         const double* __restrict__ minv,
                                                         const double* __restrict__ wcnt, double* __restrict__ z,
                                                         double* __restrict__ p, double* __restrict__ x,
                                                         double* part_rz, double* part_max, float* __restrict__ p32,
                                                         const double* __restrict__ b, const double* __restrict__ q) {
    constexpr int CTP = OpStore<float>::pad(CT);
    const FlatCols<CT> f(S);
    double rz = 0.0, mx = 0.0;
    if (f.on()) {
        if (b) {
            for (int64_t e = f.t; e < f.n; e += f.T) {
                const double rv = b[e] - q[e], mv = minv[e], n = wcnt[e];
                const double zv = mv * rv;
                r[e] = rv;
                z[e] = zv;
                rz += rv * zv;
                if (n > 0.0) mx = fmax(mx, fabs(rv) / n);
            }
        } else {
            for (int64_t e = f.t; e < f.n; e += f.T) {
                const double rv = r[e], mv = minv[e], n = wcnt[e];
                const double zv = mv * rv;
                x[e] = 0.0;
                z[e] = zv;
                p[e] = zv;
                if (p32) p32[(e / CT) * CTP + f.c] = (float)zv;
                rz += rv * zv;
                if (n > 0.0) mx = fmax(mx, fabs(rv) / n);
            }
        }
    }
    col_partials<CT>(rz, mx, part_rz, part_max);
}

// Column sums (and maxima) of per-block partials [nblk][CT], by the whole block in a fixed order;
// the results are in sum[c] / mx[c] (shared) for every thread afterwards.
template <int CT>
__device__ __forceinline__ void block_reduce_cols(const double* __restrict__ part, const double* __restrict__ part_max,
                                                  int nblk, double* sum, double* mx) {
    __shared__ double ws[2][kBlock / kWave][CT];
    double a[CT], m[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) a[c] = m[c] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            a[c] += part[b * CT + c];
            if (part_max) m[c] = fmax(m[c], part_max[b * CT + c]);
        }
    const int wv = threadIdx.x / kWave, ln = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const double sa = wave_sum(a[c]), sm = wave_max(m[c]);
        if (ln == 0) {
            ws[0][wv][c] = sa;
            ws[1][wv][c] = sm;
        }
    }
    __syncthreads();
    if (threadIdx.x < CT) {
        double ta = 0.0, tm = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) {
            ta += ws[0][w][threadIdx.x];
            tm = fmax(tm, ws[1][w][threadIdx.x]);
        }
        sum[threadIdx.x] = ta;
        mx[threadIdx.x] = tm;
    }
    __syncthreads();
}

template <int CT>
__global__ __launch_bounds__(kBlock) void cg_start_scalars(const double* part_rz, const double* part_max, int nblk,
                                                           double tol, double delta, int ncols, int first,
                                                           CgState* st) {
    __shared__ double sum[CT], mx[CT];
    block_reduce_cols<CT>(part_rz, part_max, nblk, sum, mx);
    const int c = threadIdx.x;
    if (c >= kMaxCT) return;
    if (c >= CT) {
        st->active[c] = 0;
        st->reached[c] = 1;
        return;
    }
    st->rz[c] = sum[c];
    st->maxn[c] = mx[c];
    st->tol_in[c] = fmax(tol, delta * mx[c]);
    if (first) st->iters[c] = 0;  // a restart keeps counting the steps
    st->active[c] = (c < ncols) && (mx[c] >= tol) && sum[c] > 0.0;
    st->reached[c] = !st->active[c];
    st->alpha[c] = st->beta[c] = 0.0;
}

// Reliable update: the recursive residual of each column is replaced by its true residual (cg_init_kernel
// with b, q), and the CG goes on with the same direction: beta = r'z' / r_k z_k, where r_k z_k is the
// product before the column's last step (beta and rz of that step are kept in the state), so in exact
// arithmetic the iteration is that of a CG with no update at all.  A column continues while the
// criterion of its true residual is at least tol; its next round stops at max(tol, delta x criterion).
template <int CT>
__global__ __launch_bounds__(kBlock) void cg_reliable_scalars(const double* part_rz, const double* part_max, int nblk,
                                                              double tol, double delta, int ncols, CgState* st) {
    __shared__ double sum[CT], mx[CT];
    block_reduce_cols<CT>(part_rz, part_max, nblk, sum, mx);
    const int c = threadIdx.x;
    if (c >= CT) return;
    const double rz = sum[c];
    st->beta[c] = st->rz[c] > 0.0 ? st->beta[c] * (rz / st->rz[c]) : 0.0;
    st->rz[c] = rz;
    st->maxn[c] = mx[c];
    st->tol_in[c] = fmax(tol, delta * mx[c]);
    st->active[c] = (c < ncols) && (mx[c] >= tol) && rz > 0.0;
    st->reached[c] = !st->active[c];
}

// the FP64 operator takes over: each column's rounds stop at tol itself
template <int CT>
__global__ void cg_reliable_tol(double tol, CgState* st) {
    const int c = threadIdx.x;
    if (c < CT) st->tol_in[c] = tol;
}

template <int CT>
__global__ __launch_bounds__(kBlock) void cg_pq_kernel(int64_t S, const double* __restrict__ p,
                                                       const double* __restrict__ q, double* part,
                                                       const int* __restrict__ active) {
    if (!any_active<CT>(active)) return;  // every column converged: alpha is 0 whatever the partials
    const FlatCols<CT> f(S);
    double a = 0.0;
    if (f.on())
        for (int64_t e = f.t; e < f.n; e += f.T) a += p[e] * q[e];
    col_partials<CT>(a, 0.0, part, nullptr);
}

template <int CT>
__global__ __launch_bounds__(kBlock) void cg_alpha_scalars(const double* part, int nblk, CgState* st) {
    __shared__ double sum[CT], mx[CT];
    block_reduce_cols<CT>(part, nullptr, nblk, sum, mx);
    const int c = threadIdx.x;
    if (c >= CT) return;
    const double pq = sum[c];
    st->alpha[c] = (st->active[c] && pq > 0.0) ? st->rz[c] / pq : 0.0;
}

template <int CT>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(int64_t S, const CgState* __restrict__ st,
                                                           const double* __restrict__ p, const double* __restrict__ q,
                                                           const double* __restrict__ minv,
                                                           const double* __restrict__ wcnt, double* __restrict__ x,
                                                           double* __restrict__ r, double* __restrict__ z,
                                                           double* part_rz, double* part_max) {
    // every column converged before this step: x, r and z stay as they are
    if (!any_active<CT>(st->active)) return;
    const FlatCols<CT> f(S);
    double rz = 0.0, mx = 0.0;
    if (f.on()) {
        const double al = st->alpha[f.c];
        for (int64_t e = f.t; e < f.n; e += f.T) {
            double rv = r[e];
            const double mv = minv[e], n = wcnt[e];
            if (al != 0.0) {
                const double xv = x[e], pv = p[e], qv = q[e];
                x[e] = fma(al, pv, xv);
                rv = fma(-al, qv, rv);
                r[e] = rv;
            }
            const double zv = mv * rv;
            z[e] = zv;
            rz += rv * zv;
            if (n > 0.0) mx = fmax(mx, fabs(rv) / n);
        }
    }
    col_partials<CT>(rz, mx, part_rz, part_max);
}

template <int CT>
__global__ __launch_bounds__(kBlock) void cg_beta_scalars(const double* part_rz, const double* part_max, int nblk,
                                                          double tol, CgState* st) {
    __shared__ double sum[CT], mxs[CT];
    block_reduce_cols<CT>(part_rz, part_max, nblk, sum, mxs);
    const int c = threadIdx.x;
    if (c >= CT) return;
    if (!st->active[c]) {
        st->beta[c] = 0.0;
        return;
    }
    const double rz = sum[c], mx = mxs[c];
    st->iters[c] += 1;
    st->maxn[c] = mx;
    st->beta[c] = st->rz[c] > 0.0 ? rz / st->rz[c] : 0.0;
    st->rz[c] = rz;
    if (mx < st->tol_in[c]) st->reached[c] = 1;
    if (mx < tol || !(rz > 0.0)) {
        st->active[c] = 0;
        st->reached[c] = 1;
    }
}

template <int CT>
__global__ void cg_pdir_kernel(int64_t S, const CgState* __restrict__ st, const double* __restrict__ z,
                               double* __restrict__ p, float* __restrict__ p32) {
    if (!any_active<CT>(st->active)) return;
    constexpr int CTP = OpStore<float>::pad(CT);
    const FlatCols<CT> f(S);
    if (!f.on() || !st->active[f.c]) return;
    const double be = st->beta[f.c];
    for (int64_t e = f.t; e < f.n; e += f.T) {
        const double pv = z[e] + be * p[e];
        p[e] = pv;
        if (p32) p32[(e / CT) * CTP + f.c] = (float)pv;
    }
}

// ---- FP32 rounds of the mixed-precision solve ----
// Within a round of the FP32 operator the direction lives only in its FP32 copy p32 (the operand the operator
// reads; rows padded as OpStore<float>), the recursive residual in FP32 (r32, [S][CT]) and x in FP64, and
// z = M^-1 r is formed where it is used instead of being stored.  The tile's columns share one row slot, so
// the Jacobi diagonal and the group counts are read once per table row from the slot's Schur diagonal
// (dgs, [S][2]).  A round needs its residual to fall by 1e-4 only (its rounding is below the FP32 operator's),
// and each round starts from the FP64 true residual.  Bytes per table entry and step: 64 against 116 for the
// FP64 vectors.
__device__ __forceinline__ void jacobi_entry(const double* __restrict__ dgs, int64_t g, double& minv, double& n) {
    const double d = dgs[2 * g];
    n = dgs[2 * g + 1];
    minv = (n > 0.0 && d > 1e-9 * n) ? 1.0 / d : 0.0;  // as make_minv_kernel
}

template <int CT>
__global__ __launch_bounds__(kBlock) void cg32_pq_kernel(int64_t S, const float* __restrict__ p32,
                                                         const double* __restrict__ q, double* part,
                                                         const int* __restrict__ active) {
    if (!any_active<CT>(active)) return;
    constexpr int CTP = OpStore<float>::pad(CT);
    const FlatCols<CT> f(S);
    double a = 0.0;
    if (f.on())
        for (int64_t e = f.t; e < f.n; e += f.T) a += (double)p32[(e / CT) * CTP + f.c] * q[e];
    col_partials<CT>(a, 0.0, part, nullptr);
}

// x += alpha p, r -= alpha q (r in FP32, the update formed in FP64), and the partials of r'z and of
// max |r| / n with z = M^-1 r
template <int CT>
__global__ __launch_bounds__(kBlock) void cg32_update_kernel(int64_t S, const CgState* __restrict__ st,
                                                             const float* __restrict__ p32,
                                                             const double* __restrict__ q,
                                                             const double* __restrict__ dgs, double* __restrict__ x,
                                                             float* __restrict__ r32, double* part_rz,
                                                             double* part_max) {
    if (!any_active<CT>(st->active)) return;
    constexpr int CTP = OpStore<float>::pad(CT);
    const FlatCols<CT> f(S);
    double rz = 0.0, mx = 0.0;
    if (f.on()) {
        const double al = st->alpha[f.c];
        for (int64_t e = f.t; e < f.n; e += f.T) {
            const int64_t g = e / CT;
            double minv, n;
            jacobi_entry(dgs, g, minv, n);
            double rv = (double)r32[e];
            if (al != 0.0) {
                const double xv = x[e], pv = (double)p32[g * CTP + f.c], qv = q[e];
                x[e] = fma(al, pv, xv);
                const float rn = (float)fma(-al, qv, rv);
                r32[e] = rn;
                rv = (double)rn;
            }
            rz += rv * (minv * rv);
            if (n > 0.0) mx = fmax(mx, fabs(rv) / n);
        }
    }
    col_partials<CT>(rz, mx, part_rz, part_max);
}

// p = M^-1 r + beta p on the FP32 copy
template <int CT>
__global__ void cg32_pdir_kernel(int64_t S, const CgState* __restrict__ st, const double* __restrict__ dgs,
                                 const float* __restrict__ r32, float* __restrict__ p32) {
    if (!any_active<CT>(st->active)) return;
    constexpr int CTP = OpStore<float>::pad(CT);
    const FlatCols<CT> f(S);
    if (!f.on() || !st->active[f.c]) return;
    const double be = st->beta[f.c];
    for (int64_t e = f.t; e < f.n; e += f.T) {
        const int64_t g = e / CT;
        double minv, n;
        jacobi_entry(dgs, g, minv, n);
        float* pe = p32 + g * CTP + f.c;
        *pe = (float)fma(be, (double)*pe, minv * (double)r32[e]);
    }
}

// Start of an FP32 round from the FP64 residual r and z = M^-1 r (the first start, or a reliable update):
// r32 = r, and the direction p = z + beta p on the FP32 copy (beta 0 at the first start).  to64: the solve
// leaves the FP32 operator instead, and the direction is written to the FP64 array p.
template <int CT>
__global__ void cg32_restart_kernel(int64_t S, const CgState* __restrict__ st, const double* __restrict__ r,
                                    const double* __restrict__ z, float* __restrict__ r32, float* __restrict__ p32,
                                    double* __restrict__ p, int to64) {
    constexpr int CTP = OpStore<float>::pad(CT);
    const FlatCols<CT> f(S);
    if (!f.on()) return;
    const double be = st->beta[f.c];
    const int act = st->active[f.c];
    for (int64_t e = f.t; e < f.n; e += f.T) {
        if (!to64) r32[e] = (float)r[e];
        if (act) {
            float* pe = p32 + (e / CT) * CTP + f.c;
            const double pv = fma(be, (double)*pe, z[e]);
            if (to64)
                p[e] = pv;
            else
                *pe = (float)pv;
        }
    }
}

// Per-column preconditioner and weighted counts from the per-slot Schur diagonals ([slot][S][2]).
struct ColSlots {
    int slot[kMaxCT];
};

template <int CT>
__global__ void make_minv_kernel(int64_t S, const double* __restrict__ dg, ColSlots cs, double* minv, double* wcnt) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < S; g += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const double* e = dg + ((int64_t)cs.slot[c] * S + g) * 2;
            const double d = e[0], n = e[1];
            minv[g * CT + c] = (n > 0.0 && d > 1e-9 * n) ? 1.0 / d : 0.0;
            wcnt[g * CT + c] = n;
        }
    }
}

// out[g] = tab[g][c] of a [S][CT] table
__global__ void take_column_kernel(const double* __restrict__ tab, int64_t S, int CT, int c, double* out) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < S; g += (int64_t)gridDim.x * blockDim.x)
        out[g] = tab[g * CT + c];
}

// Table blocks moved between two layouts of table space: entry g of dim block o goes from src_off[o] + g
// (zero where src_off[o] < 0 or there is no source) to dst_off[o] + g.
struct TableMap {
    int nd;
    int64_t dst_off[kMaxScatter];
    int64_t src_off[kMaxScatter];
    int G[kMaxScatter];
};

__device__ __forceinline__ int64_t map_entry(const TableMap& m, int64_t e) {
#pragma unroll
    for (int o = 0; o < kMaxScatter; ++o)
        if (o < m.nd && e >= m.dst_off[o] && e < m.dst_off[o] + m.G[o])
            return m.src_off[o] < 0 ? -1 : m.src_off[o] + (e - m.dst_off[o]);
    return -1;
}

// r[e][c] = b[map(e)][col[c]] for the tile's columns (zero past ncols): a solve's right-hand sides taken
// from a computation shared with other FE sets
struct RhsCols {
    int col[kMaxCT];
};
__global__ void rhs_gather_kernel(double* __restrict__ r, int CT, int64_t S, int ncols, const double* __restrict__ b,
                                  int ldb, TableMap m, RhsCols rc) {
    const int64_t total = S * CT;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i / CT;
        const int c = (int)(i - e * CT);
        const int64_t se = map_entry(m, e);
        r[i] = (c < ncols && se >= 0) ? b[se * ldb + rc.col[c]] : 0.0;
    }
}

__global__ void mask_words_kernel(const uint8_t* __restrict__ w, const int* __restrict__ perm, int64_t n, uint32_t bit,
                                  int first, uint32_t* wm) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const bool on = w ? w[perm ? perm[i] : i] != 0 : true;
        const uint32_t prev = first ? 0u : wm[i];
        wm[i] = prev | (on ? bit : 0u);
    }
}

// ------------------------------------------------------------------------------------------------
// host side

struct Layout {
    int nd = 0;
    int elim = 0;
    int dims[kMaxScatter];      // panel dims scattered
    int64_t S = 0;              // table groups
    int Gs = 0;                 // small groups in LDS
    ScatterDim sd[kMaxScatter];
};

int make_layout(Panel* p, unsigned fe_mask, Layout& L) {
    L = Layout();
    L.elim = (fe_mask >> p->P) & 1u;
    int large = 0;
    for (int d = 0; d < p->m; ++d) {
        if (!((fe_mask >> d) & 1u) || d == p->P) continue;
        if (L.nd == kMaxScatter) return set_error("too many FE dims");
        ScatterDim& s = L.sd[L.nd];
        s.ids = p->d[d].ids;
        s.G = p->d[d].G;
        s.small = p->d[d].small;
        s.tab_off = L.S;
        s.runbuf = nullptr;
        s.seg_run_off = nullptr;
        s.run_dest = nullptr;
        s.row_rdest = p->d[d].ids;  // a readable stand-in for dims without runs (never used as a slot)
        L.S += s.G;
        if (s.small) {
            s.lds_off = L.Gs;
            L.Gs += s.G;
            s.large_slot = -1;
        } else {
            if (large == kMaxLarge) return set_error("more than %d large non-primary FE dims", kMaxLarge);
            s.large_slot = large++;
            s.seg_run_off = p->d[d].seg_run_off;
            s.run_dest = p->d[d].run_dest;
            s.row_rdest = p->d[d].row_rdest;
            s.lds_off = 0;
        }
        L.dims[L.nd++] = d;
    }
    return 0;
}

// dynamic LDS of the tile and long-segment kernels: small-dim accumulators and staged tables
template <int CT>
size_t tables_lds(const Layout& L) {
    return sizeof(double) * (size_t)L.Gs * CT * 2;
}

unsigned capped_grid(int64_t work) {
    const int64_t cap = (int64_t)num_cus() * 8;
    return (unsigned)(work <= 0 ? 0 : (work < cap ? work : cap));
}

// all non-long segments in block tiles (Schur diagonal)
unsigned tile_grid(Panel* p) {
    const unsigned g = capped_grid(p->n_tiles);
    return g < 1 ? 1 : g;
}

// operator kernels: wave tiles (4 per block) and block tiles of medium segments
unsigned wave_grid(Panel* p) { return capped_grid((p->n_wtiles + kBlock / kWave - 1) / (kBlock / kWave)); }
unsigned mtile_grid(Panel* p) { return capped_grid(p->n_mtiles); }

unsigned long_grid(Panel* p) { return p->n_long > 0 ? stream_grid((size_t)p->n_long, kBlock, 2) : 0; }

unsigned apply_grid(Panel* p) {
    return std::max(tile_grid(p) + long_grid(p), wave_grid(p) + mtile_grid(p) + long_grid(p));
}

// One pass of the operator kernels (kIter, kData or kFinal) over every row: wave tiles, block tiles,
// long segments; returns the number of per-block partial tables written.
template <int CT, int MODE>
unsigned launch_pass(Panel* p, const Layout& L, ApplyArgs& a, hipStream_t s);

// table-space results of one apply launch: small dims from the per-block partials, large dims from
// the run partials (only_dim >= 0 restricts to one scatter dim)
template <int CT>
void reduce_tables(Panel* p, const Layout& L, const ApplyArgs& a, int only_dim, double* q, unsigned grid,
                   const int* active, hipStream_t s, bool time_runs, double on = 0.0, int run_stride = CT) {
    for (int o = 0; o < L.nd; ++o) {
        if (only_dim >= 0 && o != only_dim) continue;
        const ScatterDim& d = L.sd[o];
        double* qo = q + d.tab_off * CT;
        if (d.small) {
            hipLaunchKernelGGL((reduce_small_kernel<CT>), dim3((unsigned)(d.G * CT)), dim3(kBlock), 0, s,
                               (const double*)a.small_part, (int)grid, L.Gs, d.lds_off, d.G, qo);
        } else {
            const DimStruct& D = p->d[L.dims[o]];
            if (time_runs) kstat_begin(kStatReduceRuns, s);
            if (a.op32)
                hipLaunchKernelGGL((reduce_runs32_kernel<CT>), dim3(stream_grid((size_t)d.G * CT, kBlock, 32)),
                                   dim3(kBlock), 0, s, (const float*)a.sd[o].runbuf32, run_stride,
                                   (const int64_t*)D.grp_run_off, d.G, qo, active);
            else
                hipLaunchKernelGGL((reduce_runs_kernel<CT>), dim3(stream_grid((size_t)d.G * CT, kBlock, 32)),
                                   dim3(kBlock), 0, s, (const double*)a.sd[o].runbuf, run_stride,
                                   (const int64_t*)D.grp_run_off, d.G, qo, active);
            if (time_runs)
                kstat_end(kStatReduceRuns, s,
                          on > 0.0 ? (a.op32 ? 4.0 : 8.0) * on * (double)D.n_runs + 8.0 * on * (double)d.G +
                                         8.0 * (double)(d.G + 1)
                                   : 0.0);
        }
    }
}

// One operator product (kIter) or right-hand side (kData) into q (table space [S][CT]).
template <int CT, int MODE>
unsigned launch_pass(Panel* p, const Layout& L, ApplyArgs& a, hipStream_t s) {
    unsigned gw = wave_grid(p);
    const unsigned gm = mtile_grid(p), gl = long_grid(p);
    size_t lt = tables_lds<CT>(L);
    if (MODE == kFinal && a.gram) lt += sizeof(double) * (size_t)(kBlock / kWave) * gram_stage<CT>();
    a.part_base = 0;
    bool lean_final = false;
    if constexpr (MODE == kFinal) {
        // the lean final pass: one slot, panel-order outputs of one type, at most 3 scattered dims, and every
        // row and table offset within 32 bits
        const bool in32 = a.in32 != 0;
        const bool out32 = a.out32 == 1;
        lean_final = gw > 0 && a.lean && a.nd <= 3 && (out32 || (a.out32 == 0 && a.to_work)) &&
                     (out32 || !in32) && (uint64_t)p->n * 8u < (1ull << 32) &&
                     (uint64_t)L.S * CT * 8u < (1ull << 32);
        if (lean_final) {
            FinalArgs f{};
            f.wtile_rec = a.wtile_rec;
            f.n_wtiles = a.n_wtiles;
            f.idP = a.idP;
            f.perm = a.perm;
            f.wm = a.wm;
            f.ubit = a.ubit;
            f.elim = a.elim;
            f.ncols = a.ncols;
            for (int o = 0; o < a.nd; ++o) {
                f.ids[o] = a.sd[o].ids;
                f.tab[o] = a.alpha + a.sd[o].tab_off * CT;
            }
            for (int c = 0; c < kMaxCT; ++c) {
                const bool in = c < a.ncols;
                f.x[c] = in ? (in32 ? (const void*)a.xc32[c] : (const void*)a.xc[c]) : nullptr;
                f.out[c] = in ? (out32 ? (void*)a.oc32[c] : (void*)a.oc[c]) : nullptr;
            }
            f.gram = a.gram;
            const size_t lf = a.gram ? sizeof(double) * (size_t)(kBlock / kWave) * gram_stage<CT>() : 0;
#define FEHIP_FINAL_LAUNCH(NDV, XTV, OTV)                                                                         \
    do {                                                                                                         \
        gw = std::min(gw, 2u * resident_grid(wave_final_kernel<CT, NDV, XTV, OTV>, kBlock, lf, gw));             \
        hipLaunchKernelGGL((wave_final_kernel<CT, NDV, XTV, OTV>), dim3(gw), dim3(kBlock), lf, s, f);            \
    } while (0)
#define FEHIP_FINAL_TYPES(NDV)                                                                                    \
    do {                                                                                                         \
        if (in32) FEHIP_FINAL_LAUNCH(NDV, float, float);                                                         \
        else if (out32) FEHIP_FINAL_LAUNCH(NDV, double, float);                                                  \
        else FEHIP_FINAL_LAUNCH(NDV, double, double);                                                            \
    } while (0)
            if (a.nd == 0) FEHIP_FINAL_TYPES(0);
            else if (a.nd == 1) FEHIP_FINAL_TYPES(1);
            else if (a.nd == 2) FEHIP_FINAL_TYPES(2);
            else FEHIP_FINAL_TYPES(3);
#undef FEHIP_FINAL_TYPES
#undef FEHIP_FINAL_LAUNCH
        }
    }
    if (lean_final) {
        // wave tiles done above
    } else if (MODE == kData && a.lean && gw > 0 && !a.perm && (uint64_t)p->n * 8u < (1ull << 32)) {
        // the right-hand side through the lean kernel: one slot, rows in the caller's order
        const size_t lo = sizeof(double) * (size_t)L.Gs * CT;
#define FEHIP_DATA_LAUNCH(NDV, XTV)                                                                               \
    do {                                                                                                         \
        gw = std::min(gw, 2u * resident_grid(wave_op_kernel<CT, NDV, double, true, XTV>, kBlock, lo, gw));       \
        hipLaunchKernelGGL((wave_op_kernel<CT, NDV, double, true, XTV>), dim3(gw), dim3(kBlock), lo, s, a);      \
    } while (0)
        if (a.in32) {
            if (a.nd <= 2) FEHIP_DATA_LAUNCH(2, float);
            else if (a.nd == 3) FEHIP_DATA_LAUNCH(3, float);
            else FEHIP_DATA_LAUNCH(kMaxScatter, float);
        } else {
            if (a.nd <= 2) FEHIP_DATA_LAUNCH(2, double);
            else if (a.nd == 3) FEHIP_DATA_LAUNCH(3, double);
            else FEHIP_DATA_LAUNCH(kMaxScatter, double);
        }
#undef FEHIP_DATA_LAUNCH
    } else if (MODE == kIter && a.lean && gw > 0) {
        // the lean operator kernel: FP64 or FP32 storage, by the scattered dims at compile time
        const size_t lo = sizeof(double) * (size_t)L.Gs * CT;
#define FEHIP_OP_LAUNCH(NDV, TSV)                                                                                 \
    do {                                                                                                         \
        gw = std::min(gw, 2u * resident_grid(wave_op_kernel<CT, NDV, TSV>, kBlock, lo, gw));                     \
        hipLaunchKernelGGL((wave_op_kernel<CT, NDV, TSV>), dim3(gw), dim3(kBlock), lo, s, a);                    \
    } while (0)
        if (a.op32) {
            if (a.nd <= 2) FEHIP_OP_LAUNCH(2, float);
            else if (a.nd == 3) FEHIP_OP_LAUNCH(3, float);
            else FEHIP_OP_LAUNCH(kMaxScatter, float);
        } else {
            if (a.nd <= 2) FEHIP_OP_LAUNCH(2, double);
            else if (a.nd == 3) FEHIP_OP_LAUNCH(3, double);
            else FEHIP_OP_LAUNCH(kMaxScatter, double);
        }
#undef FEHIP_OP_LAUNCH
    } else if (gw > 0) {
        // two rounds of resident workgroups, so that workgroups of the second round fill the places of
        // the first round's early finishers (at most the panel's grid, which sizes the partial buffers)
#define FEHIP_TILE_LAUNCH(NDV, XTV)                                                                               \
    do {                                                                                                         \
        gw = std::min(gw, 2u * resident_grid(wave_tile_kernel<CT, MODE, NDV, XTV>, kBlock, lt, gw));             \
        hipLaunchKernelGGL((wave_tile_kernel<CT, MODE, NDV, XTV>), dim3(gw), dim3(kBlock), lt, s, a);            \
    } while (0)
        // the FP32 copies of registered input columns (right-hand side and final pass)
        bool done = false;
        if constexpr (MODE == kData || MODE == kFinal) {
            if (a.in32) {
                if (a.nd <= 2) FEHIP_TILE_LAUNCH(2, float);
                else if (a.nd == 3) FEHIP_TILE_LAUNCH(3, float);
                else FEHIP_TILE_LAUNCH(kMaxScatter, float);
                done = true;
            }
        }
        if (!done) {
            if (a.nd <= 2) FEHIP_TILE_LAUNCH(2, double);
            else if (a.nd == 3) FEHIP_TILE_LAUNCH(3, double);
            else FEHIP_TILE_LAUNCH(kMaxScatter, double);
        }
#undef FEHIP_TILE_LAUNCH
    }
    if (gm > 0) {
        ApplyArgs b = a;
        b.tile_seg = p->mtile_seg;
        b.n_tiles = p->n_mtiles;
        b.part_base = (int)gw;
        hipLaunchKernelGGL((tile_fast_kernel<CT, MODE>), dim3(gm), dim3(kBlock), lt, s, b);
    }
    if (gl > 0) {
        ApplyArgs b = a;
        b.part_base = (int)(gw + gm);
        hipLaunchKernelGGL((long_kernel<CT, MODE>), dim3(gl), dim3(kBlock), lt, s, b);
    }
    return gw + gm + gl;
}

// The run-level operator product q = S p (kIter with couplings): D's block from its runs, reduced per
// group, plus the couplings with the small groups.
template <int CT>
int run_op_apply(Panel* p, const Layout& L, const ApplyArgs& a, double* q, hipStream_t s, double on, bool timed,
                 int sid) {
    const RunCoupling& rc = *a.rc;
    const int o = rc.o_large;
    const bool f32 = a.op32 != 0;
    const int ctp = f32 ? OpStore<float>::pad(CT) : CT;
    if (rc.kind != 0) {
        // table-level products only: A p_D as the group counts times p_D (no primary absorbed), then the
        // couplings; the FP64 direction is read (these solves keep the FP64 operator)
        const int G = o >= 0 ? L.sd[o].G : 0;
        const int64_t off = o >= 0 ? L.sd[o].tab_off : 0;
        if (timed) kstat_begin(sid, s);
        if (G > 0) {
            hipLaunchKernelGGL((diag_scale_kernel<CT>), dim3(stream_grid((size_t)G * CT, kBlock)), dim3(kBlock), 0, s,
                               a.dg, a.p, off, G, q, a.active);
            FEHIP_LAUNCH_CHECK("diagonal product");
        }
        if (rc.Gs > 0) {
            const unsigned nblk = G > 0 ? std::max(1u, std::min<unsigned>(std::min(2048u, apply_grid(p)),
                                                                          (unsigned)((G + kCouplingTile - 1) / kCouplingTile)))
                                        : 0u;
            if (nblk > 0)
                hipLaunchKernelGGL((coupling_pass_kernel<CT, double>), dim3(nblk), dim3(kBlock), 0, s,
                                   (const double*)rc.B, rc.Gs, G, off, a.p, CT, (const int*)rc.small_rows, q,
                                   a.small_part, a.active);
            hipLaunchKernelGGL((coupling_final_kernel<CT, double>), dim3((unsigned)(rc.Gs * CT)), dim3(kBlock), 0, s,
                               (const double*)a.small_part, (int)nblk, (const double*)rc.C, rc.Gs, a.p, CT,
                               (const int*)rc.small_rows, q, a.active);
            FEHIP_LAUNCH_CHECK("couplings");
        }
        // algorithmic bytes: p and q of D's rows, B read once, the small rows
        if (timed)
            kstat_end(sid, s, on > 0.0 ? 8.0 * (double)G * (2.0 * on + 1.0) + 8.0 * (double)G * rc.Gs +
                                             8.0 * (double)rc.Gs * (rc.Gs + 2.0 * on)
                                       : 0.0);
        return 0;
    }
    const ScatterDim& d = L.sd[o];
    const DimStruct& D = p->d[L.dims[o]];
    const int* rlen = rc.run_len ? rc.run_len : (const int*)D.run_len;  // a subsample's own run lengths
    const int* glen = rc.grp_len ? rc.grp_len : (const int*)D.grp_len;
    if (timed) kstat_begin(sid, s);
#define FEHIP_RUN_OP(TSV, TAB, OUT)                                                                                 \
    do {                                                                                                           \
        if (D.n_rtiles > 0) {                                                                                      \
            const unsigned gw = std::min<unsigned>(                                                                \
                (unsigned)((D.n_rtiles + kBlock / kWave - 1) / (kBlock / kWave)),                                  \
                2u * resident_grid(run_wave_kernel<CT, TSV>, kBlock, 0, (size_t)1 << 30));                         \
            hipLaunchKernelGGL((run_wave_kernel<CT, TSV>), dim3(gw), dim3(kBlock), 0, s, TAB, ctp, d.tab_off,      \
                               (const int64_t*)D.rtile_rec, (const int64_t*)D.rtile_seg, D.n_rtiles,               \
                               (const int*)D.run_gid, rlen, OUT, a.active);                                        \
        }                                                                                                          \
        if (D.n_rlong > 0)                                                                                         \
            hipLaunchKernelGGL((run_long_kernel<CT, TSV>), dim3(stream_grid((size_t)D.n_rlong, kBlock)),            \
                               dim3(kBlock), 0, s, TAB, ctp, d.tab_off, (const int64_t*)D.rlong_segs, D.n_rlong,   \
                               (const int64_t*)D.seg_run_off, (const int*)D.run_gid, rlen, OUT, a.active);         \
    } while (0)
    if (f32)
        FEHIP_RUN_OP(float, a.p32, a.sd[o].runbuf32);
    else
        FEHIP_RUN_OP(double, a.p, a.sd[o].runbuf);
#undef FEHIP_RUN_OP
    FEHIP_LAUNCH_CHECK("run operator");
    if (timed) {
        // algorithmic bytes: the runs' group and length, D's table rows once, the segment means written
        const double es = f32 ? 4.0 : 8.0;
        const double R = (double)D.n_runs, GP = (double)p->d[p->P].G;
        kstat_end(sid, s, on > 0.0 ? 8.0 * R + es * on * ((double)d.G + GP) : 0.0);
    }
    {
        // A p_D per group from the segment means, the runs of each group in group order
        double* qo = q + d.tab_off * CT;
        const unsigned gr = stream_grid((size_t)d.G * CT, kBlock, 32);
        if (timed) kstat_begin(kStatReduceRuns, s);
        if (f32)
            hipLaunchKernelGGL((reduce_segmeans_kernel<CT, float>), dim3(gr), dim3(kBlock), 0, s, a.p32, ctp,
                               d.tab_off, (const float*)a.sd[o].runbuf32, (const int*)D.grp_seg, glen,
                               (const int64_t*)D.grp_run_off, d.G, qo, a.active);
        else
            hipLaunchKernelGGL((reduce_segmeans_kernel<CT, double>), dim3(gr), dim3(kBlock), 0, s, a.p, ctp,
                               d.tab_off, (const double*)a.sd[o].runbuf, (const int*)D.grp_seg, glen,
                               (const int64_t*)D.grp_run_off, d.G, qo, a.active);
        if (timed)
            kstat_end(kStatReduceRuns, s,
                      on > 0.0 ? 8.0 * (double)D.n_runs + (f32 ? 4.0 : 8.0) * on * ((double)D.n_runs + (double)d.G) +
                                     8.0 * on * (double)d.G + 8.0 * (double)(d.G + 1)
                               : 0.0);
    }
    if (rc.Gs > 0) {
        const int G = d.G;
        // partial blocks within the per-block partial buffer ([apply_grid][Gs][kMaxCT])
        // 2048 blocks (eight waves per SIMD) keep more of B's tiles in flight (c2: 1024 blocks 166 us, 2048 141 us)
        const unsigned nblk = std::max(
            1u, std::min<unsigned>(std::min(2048u, apply_grid(p)), (unsigned)((G + kCouplingTile - 1) / kCouplingTile)));
#define FEHIP_COUPLING(TSV, TAB, BV)                                                                               \
    do {                                                                                                           \
        hipLaunchKernelGGL((coupling_pass_kernel<CT, TSV>), dim3(nblk), dim3(kBlock), 0, s, BV,                    \
                           rc.Gs, G, d.tab_off, TAB, ctp, (const int*)rc.small_rows, q, a.small_part, a.active);   \
        hipLaunchKernelGGL((coupling_final_kernel<CT, TSV>), dim3((unsigned)(rc.Gs * CT)), dim3(kBlock), 0, s,     \
                           (const double*)a.small_part, (int)nblk, (const double*)rc.C, rc.Gs, TAB, ctp,           \
                           (const int*)rc.small_rows, q, a.active);                                                \
    } while (0)
        if (f32)
            FEHIP_COUPLING(float, a.p32, (const float*)rc.B32);  // B in FP32 for the FP32 operator
        else
            FEHIP_COUPLING(double, a.p, (const double*)rc.B);
#undef FEHIP_COUPLING
        FEHIP_LAUNCH_CHECK("couplings");
    }
    return 0;
}

// One operator product (kIter) or right-hand side (kData) into q (table space [S][CT]).  n_on: the
// columns this launch works on (the right-hand side: every column; an operator product: the columns
// whose CG is still running, which the caller knows only when it has polled the state just before;
// -1 = all), used for the algorithmic byte count of per-kernel timing.
template <int CT>
int apply_and_reduce(Panel* p, const Layout& L, ApplyArgs& a, int mode, double* q, hipStream_t s, int n_on = -1,
                     int stat = -1) {
    const int sid = stat >= 0 ? stat : mode == kIter ? (a.op32 ? kStatApplyIter32 : kStatApplyIter) : kStatApplyData;
    const bool timed = kstats_enabled();
    if (mode == kIter && a.rc) {
        const int e = run_op_apply<CT>(p, L, a, q, s, n_on < 0 ? (double)a.ncols : (double)n_on, timed, sid);
        if (e) return e;
        FEHIP_LAUNCH_CHECK("reduce kernels");
        return 0;
    }
    if (timed) kstat_begin(sid, s);
    const unsigned grid = mode == kIter ? launch_pass<CT, kIter>(p, L, a, s) : launch_pass<CT, kData>(p, L, a, s);
    FEHIP_LAUNCH_CHECK("operator kernels");
    const double on = n_on < 0 ? (double)a.ncols : (double)n_on;
    if (timed) {
        // algorithmic bytes: id streams (the scattered dims and the primary, which marks the segments),
        // row masks, table reads, run-partial and per-block partial writes of the columns worked on (+ the
        // data columns and the permutation for the right-hand side); a launch after every column has
        // converged returns at once and moves nothing
        const double n = (double)p->n;
        const double es = a.op32 ? 4.0 : 8.0;  // bytes of a table entry and of a run partial
        double bytes = n * 4.0 * (L.nd + 1) + (a.wm ? 4.0 * n : 0.0) + 8.0 * (double)grid * L.Gs * on;
        for (int o = 0; o < L.nd; ++o) {
            bytes += es * on * (double)L.sd[o].G;
            if (!L.sd[o].small) bytes += (es * on + 4.0) * (double)p->d[L.dims[o]].n_runs;
        }
        if (mode == kData) bytes += n * 8.0 * a.ncols + (p->perm ? 4.0 * n : 0.0);
        kstat_end(sid, s, on > 0.0 ? bytes : 0.0);
    }
    reduce_tables<CT>(p, L, a, -1, q, grid, mode == kIter ? a.active : nullptr, s, mode == kIter && timed, on);
    FEHIP_LAUNCH_CHECK("reduce kernels");
    return 0;
}

struct Scratch {
    std::vector<void*> owned;
    double* tab = nullptr;      // CG arrays
    double* dg = nullptr;       // per-slot Schur diagonals and weighted counts [slot][S][2]
    double* runbuf[kMaxLarge] = {nullptr, nullptr, nullptr};
    float* runbuf32[kMaxLarge] = {nullptr, nullptr, nullptr};  // FP32 operator: run partials
    float* p32 = nullptr;       // FP32 operator: the direction's FP32 copy [S][16]
    float* r32 = nullptr;       // FP32 operator: the recursive residual of a round [S][CT]
    double* bvec = nullptr;     // FP32 operator: the right-hand side [S][CT], kept for true residuals
    double* dump = nullptr;     // lean operator kernel: per-wave slots of the lanes that end no run
    double* small_part = nullptr;
    double* part = nullptr;     // reduction partials
    double* gram_rec = nullptr; // per-block cross-product records of the final pass
    CgState* st = nullptr;
    CgState* st_host = nullptr;  // per-thread pinned scratch
    const RunCoupling* rc = nullptr;  // the run-level operator's couplings (nullptr: row kernels)
    RunCoupling own_rc;               // a subsample's couplings and run lengths (built for this solve)
    ~Scratch() {
        for (void* v : owned) cache_free(v);
    }
    template <typename T>
    int alloc(T** ptr, size_t count) {
        int e = talloc(ptr, count);
        if (!e) owned.push_back((void*)*ptr);
        return e;
    }
};

constexpr int kCgBlocks = 2048;  // CG vector kernels: eight waves per SIMD keep the table loads in flight
// blocks of the CG vector kernels for a table of `elems` entries: about eight entries per thread, at most
// kCgBlocks (a small system's scalar kernels then sum fewer per-block partials)
inline int cg_blocks(int64_t elems) {
    const int64_t b = (elems + (int64_t)kBlock * 8 - 1) / ((int64_t)kBlock * 8);
    return (int)std::max<int64_t>(1, std::min<int64_t>(kCgBlocks, b));
}
// A round of the FP32-operator PCG stops once its residual criterion is this fraction of the round's true
// (FP64) residual: well above the FP32 operator's relative rounding (~1e-6), so each round's FP64
// residual falls by about this factor.
constexpr double kInnerRel = 1e-4;
constexpr int64_t kOp32MinTable = 4096;  // table entries below which a solve keeps the FP64 operator

ApplyArgs base_args(Panel* p, const Layout& L, Scratch& sc, const uint32_t* wm) {
    ApplyArgs a{};
    a.seg_off = p->seg_off;
    a.GP = p->d[p->P].G;
    a.elim = L.elim;
    a.wm = wm;
    a.nd = L.nd;
    for (int i = 0; i < L.nd; ++i) {
        a.sd[i] = L.sd[i];
        if (!L.sd[i].small) {
            a.sd[i].runbuf = sc.runbuf[L.sd[i].large_slot];
            a.sd[i].runbuf32 = sc.runbuf32[L.sd[i].large_slot];
        }
    }
    a.Gs = L.Gs;
    a.perm = p->perm;
    a.small_part = sc.small_part;
    a.tile_seg = p->tile_seg;
    a.n_tiles = p->n_tiles;
    a.idP = p->d[p->P].ids;
    a.wtile_seg = p->wtile_seg;
    a.wtile_rec = p->wtile_rec;
    a.n_wtiles = p->n_wtiles;
    a.long_segs = p->long_segs;
    a.n_long = p->n_long;
    a.dump = sc.dump;
    a.rc = sc.rc;
    a.dg = sc.dg;
    return a;
}

// Schur diagonal and weighted group counts of one slot into dg_slot ([S][2]).
int slot_diagonal(Panel* p, const Layout& L, Scratch& sc, const uint32_t* wm, uint32_t bit, double* dg_slot,
                  hipStream_t s) {
    for (int od = 0; od < L.nd; ++od) {
        ApplyArgs a = base_args(p, L, sc, wm);
        a.diag_dim = od;
        a.dbit = bit;
        a.anybit = bit;
        for (int c = 0; c < kMaxCT; ++c) a.cbit[c] = bit;
        const unsigned gt = tile_grid(p), gl = long_grid(p);
        const size_t lds = sizeof(double) * (size_t)L.Gs * 2 * 2;
        a.part_base = 0;
        hipLaunchKernelGGL(tile_diag_kernel, dim3(gt), dim3(kBlock), lds, s, a);
        FEHIP_LAUNCH_CHECK("diag kernel");
        if (gl > 0) {
            ApplyArgs b = a;
            b.part_base = (int)gt;
            hipLaunchKernelGGL((long_kernel<2, kDiag>), dim3(gl), dim3(kBlock), lds, s, b);
            FEHIP_LAUNCH_CHECK("long diag kernel");
        }
        reduce_tables<2>(p, L, a, od, dg_slot, gt + gl, nullptr, s, false);
        FEHIP_LAUNCH_CHECK("diag reduce");
    }
    return 0;
}

// The table-level operators apply to solves with at most kMaxCouplingGroups small groups and at most one
// large scattered dim D:
//  - the primary absorbed and D present with runs: A from the runs (kind 0);
//  - the primary not absorbed and D present: A is diagonal (kind 1);
//  - no large scattered dim: S = C (kind 2).
// A solve with one row slot (one subsample) takes kind 0, with its run lengths and couplings counted over the
// subsample and built for the solve: a subsample's couplings cost ceil(Gs / 16) row-operator products at 16
// columns and are not kept, which kinds 1 and 2 (diagonal or small systems, solved in a few steps) do not
// repay (c5 panel: 1.9 -> 3.2 ms per region-only request, against 8.3 -> 7.7 ms per {worker, firm, time}
// request of 18 steps).  Several slots keep the row kernels.  FEHIP_RUN_OPERATOR=0 keeps the row kernels.
int table_operator_kind(Panel* p, const Layout& L, const uint32_t* wm, int nslots) {
    if ((wm && nslots != 1) || L.S <= 0 || L.Gs > kMaxCouplingGroups) return -1;
    if (wm && !L.elim) return -1;
    const char* env = getenv("FEHIP_RUN_OPERATOR");
    if (env && env[0] == '0') return -1;
    int nl = 0, o_large = -1;
    for (int o = 0; o < L.nd; ++o)
        if (!L.sd[o].small) {
            ++nl;
            o_large = o;
        }
    if (nl == 0) return L.Gs > 0 && !wm ? 2 : -1;
    if (nl > 1) return -1;
    if (!L.elim) return 1;
    return p->d[L.dims[o_large]].run_gid != nullptr && p->d[L.dims[o_large]].n_runs > 0 ? 0 : -1;
}

// Run tiles of dim d (once per panel): whole primary segments packed greedily into at most 64 runs,
// segments of more runs listed apart.  Caller holds the panel's lock.
int build_run_tiles(Panel* p, int dd, hipStream_t s) {
    DimStruct& D = p->d[dd];
    if (D.n_rtiles >= 0) return 0;
    const int64_t GP = p->d[p->P].G;
    std::vector<int64_t> off((size_t)GP + 1);
    FEHIP_CHECK(hipMemcpyAsync(off.data(), D.seg_run_off, sizeof(int64_t) * (size_t)(GP + 1), hipMemcpyDeviceToHost, s),
                "run offsets");
    FEHIP_CHECK(hipStreamSynchronize(s), "run offsets");
    std::vector<int64_t> recs, segs, longs;
    recs.reserve((size_t)(off[GP] / 48 + 16));
    segs.reserve((size_t)(off[GP] / 48 + 16));
    int64_t w = 0;
    while (w < GP) {
        const int64_t nr = off[w + 1] - off[w];
        if (nr == 0) {
            ++w;
            continue;
        }
        if (nr > kWave) {
            longs.push_back(w);
            ++w;
            continue;
        }
        // a tile stops before an empty segment, so its segments are consecutive non-empty ones (the kernel
        // numbers them from the tile's first by their head runs)
        int64_t e = w + 1;
        while (e < GP && off[e + 1] > off[e] && off[e + 1] - off[w] <= kWave) ++e;
        recs.push_back((off[w] & 0xffffffffLL) | ((off[e] - off[w]) << 32));
        segs.push_back(w);
        w = e;
    }
    int e;
    if ((e = dalloc(&D.rtile_rec, std::max<size_t>(1, recs.size()), &p->bytes))) return e;
    if ((e = dalloc(&D.rtile_seg, std::max<size_t>(1, segs.size()), &p->bytes))) return e;
    if ((e = dalloc(&D.rlong_segs, std::max<size_t>(1, longs.size()), &p->bytes))) return e;
    if (!recs.empty()) {
        FEHIP_CHECK(hipMemcpyAsync(D.rtile_rec, recs.data(), sizeof(int64_t) * recs.size(), hipMemcpyHostToDevice, s),
                    "run tiles");
        FEHIP_CHECK(hipMemcpyAsync(D.rtile_seg, segs.data(), sizeof(int64_t) * segs.size(), hipMemcpyHostToDevice, s),
                    "run tiles");
    }
    if (!longs.empty())
        FEHIP_CHECK(hipMemcpyAsync(D.rlong_segs, longs.data(), sizeof(int64_t) * longs.size(), hipMemcpyHostToDevice, s),
                    "long run segments");
    // the runs in group order: segment and length of each
    const int64_t R = D.n_runs;
    int* run_seg = nullptr;
    if ((e = dalloc(&D.grp_seg, (size_t)std::max<int64_t>(1, R), &p->bytes)) ||
        (e = dalloc(&D.grp_len, (size_t)std::max<int64_t>(1, R), &p->bytes)) ||
        (e = dalloc(&run_seg, (size_t)std::max<int64_t>(1, R), nullptr)))
        return e;
    hipLaunchKernelGGL(run_segments_kernel, dim3(stream_grid((size_t)GP, kBlock)), dim3(kBlock), 0, s,
                       (const int64_t*)D.seg_run_off, GP, run_seg);
    hipLaunchKernelGGL(group_order_runs_kernel, dim3(stream_grid((size_t)R, kBlock)), dim3(kBlock), 0, s,
                       (const int*)D.run_order, (const int*)run_seg, (const int*)D.run_len, R, D.grp_seg, D.grp_len);
    FEHIP_LAUNCH_CHECK("group-order runs");
    FEHIP_CHECK(hipStreamSynchronize(s), "run tiles");
    dfree(run_seg);
    D.n_rlong = (int64_t)longs.size();
    D.n_rtiles = (int64_t)recs.size();
    return 0;
}

// The couplings B, C of a table-level operator: the row operator applied to unit directions of the small
// groups, sixteen at a time, gives B's and C's columns (over slot 0's rows when wm is given).  alloc(ptr,
// count) provides the device arrays.
template <typename Alloc>
int fill_couplings(Panel* p, const Layout& L, RunCoupling& rc, Scratch& sc, const uint32_t* wm, Alloc alloc,
                   hipStream_t s) {
    std::vector<int> rows;
    for (int o = 0; o < L.nd; ++o) {
        if (!L.sd[o].small) {
            rc.o_large = o;
            continue;
        }
        for (int g = 0; g < L.sd[o].G; ++g) rows.push_back((int)(L.sd[o].tab_off + g));
    }
    rc.Gs = (int)rows.size();
    const int G = rc.o_large >= 0 ? L.sd[rc.o_large].G : 0;
    const int64_t off_large = rc.o_large >= 0 ? L.sd[rc.o_large].tab_off : 0;
    const int64_t S = L.S;
    int e;
    if ((e = alloc(&rc.small_rows, (size_t)std::max(1, rc.Gs))) ||
        (e = alloc(&rc.B, (size_t)std::max<int64_t>(1, (int64_t)G * rc.Gs))) ||
        (e = alloc(&rc.B32, (size_t)std::max<int64_t>(1, (int64_t)G * rc.Gs))) ||
        (e = alloc(&rc.C, (size_t)std::max(1, rc.Gs * rc.Gs))))
        return e;
    if (rc.Gs == 0) return 0;
    FEHIP_CHECK(hipMemcpyAsync(rc.small_rows, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice, s),
                "coupling rows");
    const uint32_t bit = wm ? 1u : ~0u;
    double* P = sc.tab;               // the CG arrays are not in use yet: [S][16] directions
    double* Q = sc.tab + S * kMaxCT;  // and their products
    for (int t0 = 0; t0 < rc.Gs; t0 += kMaxCT) {
        const int nj = std::min(kMaxCT, rc.Gs - t0);
        FEHIP_CHECK(hipMemsetAsync(P, 0, sizeof(double) * (size_t)(S * kMaxCT), s), "unit directions");
        hipLaunchKernelGGL(coupling_units_kernel, dim3(1), dim3(kMaxCT), 0, s, P, kMaxCT, (const int*)rc.small_rows,
                           t0, nj);
        ApplyArgs a = base_args(p, L, sc, wm);
        a.rc = nullptr;
        a.p = P;
        a.lean = 1;
        a.ubit = bit;
        a.anybit = bit;
        for (int c = 0; c < kMaxCT; ++c) a.cbit[c] = bit;
        a.ncols = kMaxCT;
        a.active = nullptr;
        if ((e = apply_and_reduce<kMaxCT>(p, L, a, kIter, Q, s, nj, kStatCouplings))) return e;
        hipLaunchKernelGGL(coupling_extract_kernel, dim3(stream_grid((size_t)(G + rc.Gs) * nj, kBlock)),
                           dim3(kBlock), 0, s, (const double*)Q, kMaxCT, t0, nj, off_large, G, rc.Gs,
                           (const int*)rc.small_rows, rc.B, rc.C);
        FEHIP_LAUNCH_CHECK("coupling columns");
    }
    if (G > 0) {
        hipLaunchKernelGGL(to_f32_kernel, dim3(stream_grid((size_t)G * rc.Gs, kBlock)), dim3(kBlock), 0, s,
                           (const double*)rc.B, (int64_t)G * rc.Gs, rc.B32);
        FEHIP_LAUNCH_CHECK("coupling copy");
    }
    return 0;
}

// The table-level operator's couplings for this solve.  Unweighted: built on first use per absorbed-dims
// mask and kept with the panel.  One row slot (a subsample): built for the solve, with the large dim's run
// lengths counted over the subsample (kind 0).  Runs on the calling thread's stream before any capture.
int run_couplings(Panel* p, const Layout& L, unsigned mask, int kind, Scratch& sc, const uint32_t* wm,
                  hipStream_t s) {
    std::lock_guard<std::recursive_mutex> guard(p->mu);
    int e;
    if (kind == 0)
        for (int o = 0; o < L.nd; ++o)
            if (!L.sd[o].small && (e = build_run_tiles(p, L.dims[o], s))) return e;
    if (wm) {
        RunCoupling& rc = sc.own_rc;
        rc = RunCoupling();
        rc.kind = kind;
        if ((e = fill_couplings(p, L, rc, sc, wm, [&](auto** ptr, size_t n) { return sc.alloc(ptr, n); }, s)))
            return e;
        if (kind == 0) {
            const DimStruct& D = p->d[L.dims[rc.o_large]];
            const int64_t R = D.n_runs;
            int* lw = nullptr;
            int* glw = nullptr;
            if ((e = sc.alloc(&lw, (size_t)std::max<int64_t>(1, R))) || (e = sc.alloc(&glw, (size_t)std::max<int64_t>(1, R))))
                return e;
            if (p->n_mtiles == 0 && p->n_long == 0 && p->n_wtiles > 0)  // every segment in a wave tile
                hipLaunchKernelGGL(weighted_runs_wave_kernel, dim3(stream_grid((size_t)p->n_wtiles, kBlock / kWave)),
                                   dim3(kBlock), 0, s, (const int64_t*)p->wtile_rec, p->n_wtiles,
                                   (const int*)D.row_rdest, wm, 1u, (const int*)D.run_order, (const int*)D.run_len,
                                   glw, lw);
            else
                hipLaunchKernelGGL(weighted_runs_kernel, dim3(stream_grid((size_t)p->n, kBlock)), dim3(kBlock), 0, s,
                                   (const int*)p->d[p->P].ids, (const int*)D.ids, (const int*)D.row_rdest, wm, 1u,
                                   (const int*)D.run_order, (const int*)D.run_len, p->n, glw, lw);
            FEHIP_LAUNCH_CHECK("subsample run lengths");
            rc.run_len = lw;
            rc.grp_len = glw;
        }
        sc.rc = &rc;
        return 0;
    }
    auto it = p->run_couplings.find(mask);
    if (it != p->run_couplings.end()) {
        sc.rc = &it->second;
        return 0;
    }
    RunCoupling rc;
    rc.kind = kind;
    if ((e = fill_couplings(p, L, rc, sc, nullptr, [&](auto** ptr, size_t n) { return dalloc(ptr, n, &p->bytes); },
                            s)))
        return e;
    FEHIP_CHECK(hipStreamSynchronize(s), "couplings");
    auto res = p->run_couplings.emplace(mask, rc);
    sc.rc = &res.first->second;
    return 0;
}

// Per-thread stream for capturing CG steps.  Its work is ordered after the thread's default stream by
// an event before the loop, and the host waits for it after the loop.
// Graph capture of the CG step is per thread: while a capture is open the runtime rejects blocking
// calls (hipMemcpy, hipMalloc, ...) from every thread of the process, so threads that run alongside
// other device work (the request server's workers) solve with direct launches.
thread_local bool t_capture = true;

hipStream_t graph_stream() {
    struct Holder {
        hipStream_t s = nullptr;
        ~Holder() {
            if (s) (void)hipStreamDestroy(s);
        }
    };
    thread_local Holder h;
    if (!h.s && hipStreamCreateWithFlags(&h.s, hipStreamNonBlocking) != hipSuccess) h.s = nullptr;
    return h.s;
}

template <int CT>
int solve_tile(Panel* p, const Layout& L, Scratch& sc, const WithinCols& wc, int c0, int ncols, const uint32_t* wm,
               const fehip_demean_opts& o, int to_work, TileStats& ts, hipStream_t s) {
    const int64_t S = L.S;
    const int nbcg = cg_blocks(S * CT);  // blocks of the CG vector kernels
    ApplyArgs a = base_args(p, L, sc, wm);
    ColSlots cs{};
    a.anybit = 0;
    for (int c = 0; c < kMaxCT; ++c) {
        const bool in = c < ncols;
        const int slot = wc.slot[c0 + (in ? c : 0)];  // padding columns borrow the first column's slot
        a.cbit[c] = in ? (wm ? (1u << slot) : ~0u) : 0u;
        a.anybit |= a.cbit[c];
        a.xc[c] = in ? wc.x[c0 + c] : nullptr;
        a.xc32[c] = (in && !wc.x32.empty()) ? wc.x32[c0 + c] : nullptr;
        a.oc[c] = in ? wc.out[c0 + c] : nullptr;
        a.oc32[c] = in ? wc.out32[c0 + c] : nullptr;
        cs.slot[c] = slot;
    }
    // the right-hand side and the final pass read the FP32 copies of registered columns when every column
    // of the tile has one and every primary segment lies in a wave tile (the block and long kernels read FP64)
    a.in32 = p->n_mtiles == 0 && p->n_long == 0 && ncols > 0 ? 1 : 0;
    for (int c = 0; c < ncols; ++c) a.in32 = a.in32 && a.xc32[c] != nullptr;
    a.ncols = ncols;
    a.to_work = to_work;
    int n32 = 0;
    for (int c = 0; c < ncols; ++c) n32 += a.oc32[c] != nullptr ? 1 : 0;
    a.out32 = n32 == 0 ? 0 : (n32 == ncols ? 1 : 2);
    double* xa = sc.tab;           // alpha
    double* r = xa + S * CT;
    double* z = r + S * CT;
    double* pp = z + S * CT;
    double* q = pp + S * CT;
    double* minv = q + S * CT;     // [S][CT]
    double* wcnt = minv + S * CT;  // [S][CT]
    auto t0 = std::chrono::steady_clock::now();
    ts = TileStats();
    for (int c = 0; c < ncols; ++c) ts.col_converged[c] = 1;
    // one slot for every column of the tile: the lean operator kernel
    bool uniform = true;
    for (int c = 1; c < ncols; ++c) uniform = uniform && wc.slot[c0 + c] == wc.slot[c0];
    a.lean = uniform ? 1 : 0;
    a.ubit = wm ? (1u << wc.slot[c0]) : ~0u;
    // the FP32 operator (mixed precision asked, one slot, every primary segment in wave tiles): rounds of
    // FP32-operator PCG, each stopped once its own residual is delta times the round's true residual,
    // with the true residual b - S x of the FP64 operator between rounds; the stopping rule is checked
    // on that FP64 residual only, so the result meets the same criterion as the FP64 solve
    // A system that is diagonal (one scattered dim, primary not absorbed: Jacobi CG ends in one step) or
    // small (a few thousand table entries: steps are cheap) gains nothing from FP32 steps, and would pay
    // their FP64 true-residual products, so it stays on the FP64 operator.
    const bool cheap = (L.nd == 1 && !L.elim) || S <= kOp32MinTable || (sc.rc && sc.rc->kind != 0);
    const bool op32 = o.precision == 1 && uniform && sc.p32 != nullptr && !cheap;
    const double delta = op32 ? kInnerRel : 0.0;
    if (S > 0) {
        hipLaunchKernelGGL((make_minv_kernel<CT>), dim3(stream_grid(S, kBlock)), dim3(kBlock), 0, s, S,
                           (const double*)sc.dg, cs, minv, wcnt);
        FEHIP_LAUNCH_CHECK("minv");
        int e = 0;
        if (wc.rhs) {
            // b from the right-hand sides shared with other FE sets: this layout's dim blocks, the tile's columns
            TableMap m{};
            m.nd = L.nd;
            for (int o = 0; o < L.nd; ++o) {
                m.dst_off[o] = L.sd[o].tab_off;
                m.src_off[o] = wc.rhs->dim_off[L.dims[o]];
                m.G[o] = L.sd[o].G;
                if (m.src_off[o] < 0) return set_error("shared right-hand sides lack FE dim %d", L.dims[o] + 1);
            }
            RhsCols rcol{};
            for (int c = 0; c < kMaxCT; ++c) rcol.col[c] = c < ncols ? wc.rhs_col[c0 + c] : 0;
            hipLaunchKernelGGL(rhs_gather_kernel, dim3(stream_grid((size_t)(S * CT), kBlock)), dim3(kBlock), 0, s, r, CT,
                               S, ncols, wc.rhs->b, wc.rhs->ld, m, rcol);
            FEHIP_LAUNCH_CHECK("shared right-hand sides");
        } else {
            e = apply_and_reduce<CT>(p, L, a, kData, r, s);  // b = D_o' W M_P x
            if (e) return e;
        }
        if (op32) FEHIP_CHECK(hipMemcpyAsync(sc.bvec, r, sizeof(double) * (size_t)(S * CT), hipMemcpyDeviceToDevice, s),
                              "right-hand side copy");
        hipLaunchKernelGGL((cg_init_kernel<CT>), dim3(nbcg), dim3(kBlock), 0, s, S, r, (const double*)minv,
                           (const double*)wcnt, z, pp, xa, sc.part, sc.part + kCgBlocks * CT, op32 ? sc.p32 : nullptr,
                           (const double*)nullptr, (const double*)nullptr);
        hipLaunchKernelGGL((cg_start_scalars<CT>), dim3(1), dim3(kBlock), 0, s, (const double*)sc.part,
                           (const double*)sc.part + kCgBlocks * CT, nbcg, o.tol, delta, ncols, 1, sc.st);
        // the FP32 rounds' Jacobi diagonal and counts: the tile's slot's Schur diagonal
        const double* dgs = sc.dg + (int64_t)wc.slot[c0] * S * 2;
        if (op32)
            hipLaunchKernelGGL((cg32_restart_kernel<CT>), dim3(stream_grid(S, kBlock)), dim3(kBlock), 0, s, S,
                               (const CgState*)sc.st, (const double*)r, (const double*)z, sc.r32, sc.p32, pp, 0);
        FEHIP_LAUNCH_CHECK("cg init");
        const int every = o.check_every > 0 ? o.check_every : 4;
        a.p = pp;
        a.op32 = op32 ? 1 : 0;
        a.p32 = op32 ? sc.p32 : nullptr;
        a.active = sc.st->active;
        // one CG step: operator product, p.q, alpha, update, r.z, beta, new direction
        int n_on = -1;  // columns still running (known under per-kernel timing, which polls every step)
        auto step = [&](hipStream_t st) -> int {
            int e2 = apply_and_reduce<CT>(p, L, a, kIter, q, st, n_on);
            if (e2) return e2;
            if (a.op32) {
                // FP32 round: the direction's FP32 copy, the FP32 residual, x in FP64
                hipLaunchKernelGGL((cg32_pq_kernel<CT>), dim3(nbcg), dim3(kBlock), 0, st, S,
                                   (const float*)sc.p32, (const double*)q, sc.part, (const int*)sc.st->active);
                hipLaunchKernelGGL((cg_alpha_scalars<CT>), dim3(1), dim3(kBlock), 0, st, (const double*)sc.part,
                                   nbcg, sc.st);
                hipLaunchKernelGGL((cg32_update_kernel<CT>), dim3(nbcg), dim3(kBlock), 0, st, S,
                                   (const CgState*)sc.st, (const float*)sc.p32, (const double*)q, dgs, xa, sc.r32,
                                   sc.part, sc.part + kCgBlocks * CT);
                hipLaunchKernelGGL((cg_beta_scalars<CT>), dim3(1), dim3(kBlock), 0, st, (const double*)sc.part,
                                   (const double*)sc.part + kCgBlocks * CT, nbcg, o.tol, sc.st);
                hipLaunchKernelGGL((cg32_pdir_kernel<CT>), dim3(stream_grid(S, kBlock)), dim3(kBlock), 0, st, S,
                                   (const CgState*)sc.st, dgs, (const float*)sc.r32, sc.p32);
                FEHIP_LAUNCH_CHECK("cg step");
                return 0;
            }
            hipLaunchKernelGGL((cg_pq_kernel<CT>), dim3(nbcg), dim3(kBlock), 0, st, S, (const double*)pp,
                               (const double*)q, sc.part, (const int*)sc.st->active);
            hipLaunchKernelGGL((cg_alpha_scalars<CT>), dim3(1), dim3(kBlock), 0, st, (const double*)sc.part, nbcg,
                               sc.st);
            hipLaunchKernelGGL((cg_update_kernel<CT>), dim3(nbcg), dim3(kBlock), 0, st, S, (const CgState*)sc.st,
                               (const double*)pp, (const double*)q, (const double*)minv, (const double*)wcnt, xa, r, z,
                               sc.part, sc.part + kCgBlocks * CT);
            hipLaunchKernelGGL((cg_beta_scalars<CT>), dim3(1), dim3(kBlock), 0, st, (const double*)sc.part,
                               (const double*)sc.part + kCgBlocks * CT, nbcg, o.tol, sc.st);
            hipLaunchKernelGGL((cg_pdir_kernel<CT>), dim3(stream_grid(S, kBlock)), dim3(kBlock), 0, st, S,
                               (const CgState*)sc.st, (const double*)z, pp, a.op32 ? sc.p32 : nullptr);
            FEHIP_LAUNCH_CHECK("cg step");
            return 0;
        };
        // The step is captured once into a HIP graph and replayed, so an iteration costs one launch (the
        // step holds no memset and no host work).  With per-kernel timing on, the step is launched
        // directly so the kernels can be timed one by one.
        hipStream_t gs = s;
        hipGraphExec_t gexec = nullptr;
        struct GraphGuard {
            hipGraphExec_t* ex;
            ~GraphGuard() {
                if (*ex) (void)hipGraphExecDestroy(*ex);
            }
        } gguard{&gexec};
        if (t_capture && !kstats_enabled() && o.max_iter > every) {
            gs = graph_stream();
            if (!gs) return set_error("cg: no capture stream");
            // the capture stream starts after everything already queued on this thread's stream
            hipEvent_t ready;
            FEHIP_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming), "cg event");
            FEHIP_CHECK(hipEventRecord(ready, s), "cg event");
            FEHIP_CHECK(hipStreamWaitEvent(gs, ready, 0), "cg event");
            (void)hipEventDestroy(ready);
            hipGraph_t graph = nullptr;
            std::lock_guard<std::mutex> capture_lock(capture_mutex());
            FEHIP_CHECK(hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal), "cg capture");
            // one graph holds the `every` steps between two convergence checks
            int ec = 0;
            for (int k = 0; k < every && !ec; ++k) ec = step(gs);
            const hipError_t he = hipStreamEndCapture(gs, &graph);
            if (ec) {
                if (graph) (void)hipGraphDestroy(graph);
                return ec;
            }
            FEHIP_CHECK(he, "cg capture end");
            const hipError_t hi = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            FEHIP_CHECK(hi, "cg graph");
        }
        int it = 0;       // CG steps of every round
        int refines = 0;  // FP64 true-residual products (FP32 operator)
        double prev_worst = HUGE_VAL;
        for (;;) {
            int j = 0;  // steps of this round (the graph holds `every` of them)
            while (it < o.max_iter) {
                if (j % every == 0) {
                    FEHIP_CHECK(hipMemcpyAsync(sc.st_host, sc.st, sizeof(CgState), hipMemcpyDeviceToHost, gs),
                                "cg state");
                    FEHIP_CHECK(hipStreamSynchronize(gs), "cg sync");
                    bool pending = false;  // a column still short of this round's level
                    for (int c = 0; c < CT; ++c) pending |= sc.st_host->active[c] != 0 && !sc.st_host->reached[c];
                    if (!pending) break;
                }
                if (!gexec && kstats_enabled()) {
                    // per-kernel timing: the state is read before every step so the step's byte count covers
                    // the columns it works on; the launches are those of the untimed path
                    FEHIP_CHECK(hipMemcpyAsync(sc.st_host, sc.st, sizeof(CgState), hipMemcpyDeviceToHost, gs),
                                "cg state");
                    FEHIP_CHECK(hipStreamSynchronize(gs), "cg sync");
                    n_on = 0;
                    for (int c = 0; c < ncols; ++c) n_on += sc.st_host->active[c] != 0;
                }
                if (gexec && j % every == 0 && it + every <= o.max_iter) {
                    FEHIP_CHECK(hipGraphLaunch(gexec, gs), "cg graph launch");
                    it += every;
                    j += every;
                    continue;
                }
                if ((e = step(gs))) return e;  // per-kernel timing, or the steps past the last whole graph
                ++it;
                ++j;
            }
            if (!op32) break;
            // true residual r = b - S x with the FP64 operator, then a restart of the CG from it
            ApplyArgs ar = a;
            ar.p = xa;
            ar.op32 = 0;
            ar.p32 = nullptr;
            ar.active = nullptr;
            if ((e = apply_and_reduce<CT>(p, L, ar, kIter, q, gs, -1, kStatApplyResid))) return e;
            ++refines;
            const bool fp32_round = a.op32 != 0;
            hipLaunchKernelGGL((cg_init_kernel<CT>), dim3(nbcg), dim3(kBlock), 0, gs, S, r, (const double*)minv,
                               (const double*)wcnt, z, pp, xa, sc.part, sc.part + kCgBlocks * CT, (float*)nullptr,
                               (const double*)sc.bvec, (const double*)q);
            FEHIP_LAUNCH_CHECK("cg reliable update");
            // the new direction waits for the decision below: a round that did not halve the true residual
            // goes on with the FP64 operator (the FP32 operator's rounding is in the way), launched directly
            hipLaunchKernelGGL((cg_reliable_scalars<CT>), dim3(1), dim3(kBlock), 0, gs, (const double*)sc.part,
                               (const double*)sc.part + kCgBlocks * CT, nbcg, o.tol, fp32_round ? delta : 0.0,
                               ncols, sc.st);
            FEHIP_CHECK(hipMemcpyAsync(sc.st_host, sc.st, sizeof(CgState), hipMemcpyDeviceToHost, gs), "cg state");
            FEHIP_CHECK(hipStreamSynchronize(gs), "cg sync");
            bool any = false;
            double worst = 0.0;
            for (int c = 0; c < ncols; ++c)
                if (sc.st_host->active[c]) {
                    any = true;
                    worst = std::max(worst, sc.st_host->maxn[c]);
                }
            if (!any || it >= o.max_iter) break;
            if (fp32_round && !(worst < 0.5 * prev_worst)) {
                a.op32 = 0;
                a.p32 = nullptr;
                if (gexec) {
                    (void)hipGraphExecDestroy(gexec);
                    gexec = nullptr;
                }
                hipLaunchKernelGGL((cg_reliable_tol<CT>), dim3(1), dim3(kMaxCT), 0, gs, o.tol, sc.st);
            }
            // the direction after the update: the FP32 copy for another FP32 round, the FP64 direction from the
            // FP32 copy when the solve leaves the FP32 operator, and from the FP64 one after an FP64 round
            if (fp32_round)
                hipLaunchKernelGGL((cg32_restart_kernel<CT>), dim3(stream_grid(S, kBlock)), dim3(kBlock), 0, gs, S,
                                   (const CgState*)sc.st, (const double*)r, (const double*)z, sc.r32, sc.p32, pp,
                                   a.op32 ? 0 : 1);
            else
                hipLaunchKernelGGL((cg_pdir_kernel<CT>), dim3(stream_grid(S, kBlock)), dim3(kBlock), 0, gs, S,
                                   (const CgState*)sc.st, (const double*)z, pp, (float*)nullptr);
            FEHIP_LAUNCH_CHECK("cg reliable update");
            prev_worst = worst;
        }
        FEHIP_CHECK(hipMemcpyAsync(sc.st_host, sc.st, sizeof(CgState), hipMemcpyDeviceToHost, gs), "cg state");
        FEHIP_CHECK(hipStreamSynchronize(gs), "cg sync");
        for (int c = 0; c < ncols; ++c) {
            ts.iterations = std::max(ts.iterations, sc.st_host->iters[c]);
            if (sc.st_host->active[c]) ts.converged = 0;
            ts.max_update = std::max(ts.max_update, sc.st_host->maxn[c]);
            ts.col_iters[c] = sc.st_host->iters[c];
            ts.col_converged[c] = !sc.st_host->active[c];
            ts.col_max_update[c] = sc.st_host->maxn[c];
        }
        ts.apply_launches = it;
        ts.refine_launches = refines;
        for (int c = 0; c < ncols; ++c)
            if (wc.alpha[c0 + c])
                hipLaunchKernelGGL(take_column_kernel, dim3(stream_grid(S, kBlock)), dim3(kBlock), 0, s,
                                   (const double*)xa, S, CT, c, wc.alpha[c0 + c]);
        FEHIP_LAUNCH_CHECK("keep tables");
    }
    // final pass over the columns: M_P (x - D_o a), with the cross products of the tile if asked
    a.alpha = xa;
    a.gram = wc.gram ? sc.gram_rec : nullptr;
    if (kstats_enabled()) kstat_begin(kStatFinalize, s);
    const unsigned nrec = launch_pass<CT, kFinal>(p, L, a, s);
    FEHIP_LAUNCH_CHECK("finalize kernels");
    if (wc.gram && nrec > 0) {
        const unsigned nb = (nrec + kGramRun - 1) / kGramRun;
        double* gpart = sc.gram_rec + (size_t)nrec * 256;
        hipLaunchKernelGGL(gram_sum_kernel, dim3(nb), dim3(256), 0, s, (const double*)sc.gram_rec, (int)nrec, gpart);
        hipLaunchKernelGGL(gram_final_kernel, dim3(1), dim3(256), 0, s, (const double*)gpart, (int)nb, ncols,
                           wc.gram + (int64_t)c0 * wc.gram_ld + c0, wc.gram_ld);
        FEHIP_LAUNCH_CHECK("cross products");
    }
    if (kstats_enabled()) {
        double bytes = 16.0 * (double)p->n * ncols + 4.0 * (double)p->n * (L.nd + 1) +
                       (a.wm ? 4.0 * (double)p->n : 0.0) + (p->perm ? 4.0 * (double)p->n : 0.0);
        for (int od = 0; od < L.nd; ++od) bytes += 8.0 * ncols * (double)L.sd[od].G;
        kstat_end(kStatFinalize, s, bytes);
    }
    FEHIP_CHECK(hipStreamSynchronize(s), "within sync");
    ts.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

// ---- column-tile autotuner -------------------------------------------------------------------
// The operator kernels are instantiated for tile widths 1, 2, 3, 4, 6, 8, 9, 12 and 16 columns.  A
// wider tile shares the id streams and segment structure among more columns but holds more registers
// (fewer waves in flight), so the best split of a solve's columns depends on the panel.  The tuner
// times one operator product (all columns active) per width on first use of a panel and absorbed-dims
// mask, and the plan then minimises the summed product time of the tiles.
constexpr int kWidths[] = {1, 2, 3, 4, 6, 8, 9, 12, 16};
constexpr int kNumWidths = 9;

template <int CT>
int time_width(Panel* p, const Layout& L, Scratch& sc, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, float* ms) {
    ApplyArgs a = base_args(p, L, sc, nullptr);
    for (int c = 0; c < kMaxCT; ++c) {
        a.cbit[c] = c < CT ? ~0u : 0u;
        a.xc[c] = nullptr;
        a.oc[c] = nullptr;
        a.oc32[c] = nullptr;
    }
    a.anybit = ~0u;
    a.ncols = CT;
    a.p = sc.tab;
    a.active = nullptr;
    a.lean = 1;  // the kernel of an unweighted single-request solve
    a.ubit = ~0u;
    double* q = sc.tab + L.S * kMaxCT;
    auto product = [&]() -> int {
        if (a.rc) return run_op_apply<CT>(p, L, a, q, s, (double)CT, false, kStatApplyIter);
        const unsigned grid = launch_pass<CT, kIter>(p, L, a, s);
        reduce_tables<CT>(p, L, a, -1, q, grid, nullptr, s, false);
        return 0;
    };
    int e = product();  // warm-up
    if (e) return e;
    FEHIP_CHECK(hipEventRecord(ev0, s), "tuner event");
    if ((e = product())) return e;
    FEHIP_CHECK(hipEventRecord(ev1, s), "tuner event");
    FEHIP_CHECK(hipEventSynchronize(ev1), "tuner sync");
    FEHIP_CHECK(hipEventElapsedTime(ms, ev0, ev1), "tuner time");
    return 0;
}

int plan_tiles(Panel* p, const Layout& L, Scratch& sc, unsigned mask, int ncols, int ct_max, std::vector<int>& tiles,
               hipStream_t s) {
    tiles.clear();
    auto round_up = [&](int c) {
        for (int w : kWidths)
            if (w >= c) return w;
        return 16;
    };
    if (ncols <= 0) return 0;
    if (L.S == 0 || ncols == 1) {  // nothing to iterate on, or nothing to choose
        for (int c0 = 0; c0 < ncols; c0 += ct_max) tiles.push_back(round_up(std::min(ct_max, ncols - c0)));
        return 0;
    }
    std::lock_guard<std::recursive_mutex> guard(p->mu);
    auto it = p->width_ms.find(mask);
    if (it == p->width_ms.end()) {
        std::vector<float> t(kNumWidths, -1.0f);
        FEHIP_CHECK(hipMemsetAsync(sc.tab, 0, sizeof(double) * (size_t)(L.S * kMaxCT * 2), s), "tuner tables");
        hipEvent_t ev0, ev1;
        FEHIP_CHECK(hipEventCreate(&ev0), "tuner event");
        FEHIP_CHECK(hipEventCreate(&ev1), "tuner event");
        int e = 0;
        for (int k = 0; k < kNumWidths && !e; ++k) {
            if (kWidths[k] > ct_max) continue;
            float ms = 0.0f;
            switch (kWidths[k]) {
                case 1: e = time_width<1>(p, L, sc, s, ev0, ev1, &ms); break;
                case 2: e = time_width<2>(p, L, sc, s, ev0, ev1, &ms); break;
                case 3: e = time_width<3>(p, L, sc, s, ev0, ev1, &ms); break;
                case 4: e = time_width<4>(p, L, sc, s, ev0, ev1, &ms); break;
                case 6: e = time_width<6>(p, L, sc, s, ev0, ev1, &ms); break;
                case 8: e = time_width<8>(p, L, sc, s, ev0, ev1, &ms); break;
                case 9: e = time_width<9>(p, L, sc, s, ev0, ev1, &ms); break;
                case 12: e = time_width<12>(p, L, sc, s, ev0, ev1, &ms); break;
                default: e = time_width<16>(p, L, sc, s, ev0, ev1, &ms); break;
            }
            t[k] = ms;
        }
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
        if (e) return e;
        it = p->width_ms.emplace(mask, t).first;
    }
    const std::vector<float>& t = it->second;
    auto cost_of = [&](int w) {
        for (int k = 0; k < kNumWidths; ++k)
            if (kWidths[k] == w) return t[k] >= 0.0f ? (double)t[k] : 1e30;
        return 1e30;
    };
    // full tiles of width w and one tile for the remainder, rounded up to an instantiated width
    double best = 1e300;
    int best_w = round_up(std::min(ct_max, ncols));
    for (int w : kWidths) {
        if (w > ct_max) continue;
        const int full = ncols / w, rem = ncols - full * w;
        const double c = full * cost_of(w) + (rem ? cost_of(round_up(rem)) : 0.0);
        if (c < best) {
            best = c;
            best_w = w;
        }
    }
    const int full = ncols / best_w, rem = ncols - full * best_w;
    for (int k = 0; k < full; ++k) tiles.push_back(best_w);
    if (rem) tiles.push_back(round_up(rem));
    return 0;
}

}  // namespace

void set_thread_capture(bool on) { t_capture = on; }
bool thread_capture() { return t_capture; }

int build_mask_words(Panel* p, const uint8_t* const* w_caller, int nslots, uint32_t* wm, hipStream_t s) {
    if (nslots < 1 || nslots > 32) return set_error("1..32 request slots per batch (%d)", nslots);
    for (int b = 0; b < nslots; ++b) {
        hipLaunchKernelGGL(mask_words_kernel, dim3(stream_grid((size_t)p->n, kBlock)), dim3(kBlock), 0, s, w_caller[b],
                           (const int*)p->perm, p->n, 1u << b, b == 0 ? 1 : 0, wm);
        FEHIP_LAUNCH_CHECK("mask words");
    }
    return 0;
}

int64_t within_table_size(Panel* p, unsigned fe_mask) {
    unsigned mask = fe_mask ? fe_mask : ((1u << p->m) - 1u);
    mask &= (1u << p->m) - 1u;
    Layout L;
    if (make_layout(p, mask, L)) return -1;
    return L.S;
}

int within_sweep_with_tables(Panel* p, unsigned fe_mask, const double* x, const double* alpha, const uint32_t* wm,
                             int slot, double* out, hipStream_t s) {
    unsigned mask = fe_mask ? fe_mask : ((1u << p->m) - 1u);
    mask &= (1u << p->m) - 1u;
    Layout L;
    int e = make_layout(p, mask, L);
    if (e) return e;
    if ((size_t)L.Gs * 16 > 28 * 1024) return set_error("small FE dims exceed the LDS budget (%d groups)", L.Gs);
    Scratch sc;
    ApplyArgs a = base_args(p, L, sc, wm);
    for (int c = 0; c < kMaxCT; ++c) {
        a.cbit[c] = 0u;
        a.xc[c] = nullptr;
        a.oc[c] = nullptr;
        a.oc32[c] = nullptr;
    }
    a.cbit[0] = wm ? (1u << slot) : ~0u;
    a.anybit = a.cbit[0];
    a.xc[0] = x;
    a.oc[0] = out;
    a.ncols = 1;
    a.to_work = 1;
    a.alpha = alpha;
    launch_pass<1, kFinal>(p, L, a, s);
    FEHIP_LAUNCH_CHECK("sweep with tables");
    return 0;
}

// ---- right-hand sides shared by the solves of several unweighted FE sets (a batch's requests) ----

static int tile_cap(const Layout& L);

int within_dim_offsets(Panel* p, unsigned fe_mask, int64_t* off) {
    unsigned mask = fe_mask ? fe_mask : ((1u << p->m) - 1u);
    mask &= (1u << p->m) - 1u;
    Layout L;
    const int e = make_layout(p, mask, L);
    if (e) return e;
    for (int d = 0; d < FEHIP_MAX_FE; ++d) off[d] = -1;
    for (int o = 0; o < L.nd; ++o) off[L.dims[o]] = L.sd[o].tab_off;
    return 0;
}

// scratch of a right-hand-side pass: run partials of the large dims and per-block small-dim partials
static int pass_scratch(Panel* p, const Layout& L, Scratch& sc) {
    int e;
    for (int i = 0; i < L.nd; ++i)
        if (!L.sd[i].small) {
            const int64_t R = p->d[L.dims[i]].n_runs;
            if ((e = sc.alloc(&sc.runbuf[L.sd[i].large_slot], (size_t)(R * kMaxCT + 1)))) return e;
        }
    const unsigned grid = apply_grid(p);
    if ((e = sc.alloc(&sc.small_part, (size_t)grid * (L.Gs > 0 ? L.Gs : 1) * kMaxCT))) return e;
    if ((e = sc.alloc(&sc.dump, (size_t)wave_grid(p) * (kBlock / kWave) * kMaxCT + kMaxCT))) return e;
    return 0;
}

static int width_for(int n) {
    for (int w : kWidths)
        if (w >= n) return w;
    return 16;
}

template <int CT>
static int rhs_cols(Panel* p, const Layout& L, Scratch& sc, const double* const* x, const float* const* x32, int nx,
                    double* b, hipStream_t s) {
    ApplyArgs a = base_args(p, L, sc, nullptr);
    a.rc = nullptr;
    for (int c = 0; c < kMaxCT; ++c) {
        const bool in = c < nx;
        a.cbit[c] = in ? ~0u : 0u;
        a.xc[c] = in ? x[c] : nullptr;
        a.xc32[c] = (in && x32) ? x32[c] : nullptr;
        a.oc[c] = nullptr;
        a.oc32[c] = nullptr;
    }
    a.anybit = ~0u;
    a.ubit = ~0u;
    a.lean = 1;  // one slot: every row of every column
    a.ncols = nx;
    a.in32 = x32 && p->n_mtiles == 0 && p->n_long == 0 ? 1 : 0;
    for (int c = 0; c < nx; ++c) a.in32 = a.in32 && a.xc32[c] != nullptr;
    return apply_and_reduce<CT>(p, L, a, kData, b, s);
}

int within_rhs(Panel* p, unsigned fe_mask, const double* const* x, const float* const* x32, int nx, double* b, int* ld,
               hipStream_t s) {
    unsigned mask = fe_mask ? fe_mask : ((1u << p->m) - 1u);
    mask &= (1u << p->m) - 1u;
    Layout L;
    int e = make_layout(p, mask, L);
    if (e) return e;
    if (nx < 1 || nx > kMaxCT) return set_error("within_rhs: 1..%d columns (%d)", kMaxCT, nx);
    const int ct = width_for(nx);
    if (ct > tile_cap(L)) return set_error("within_rhs: small FE dims exceed the LDS budget at %d columns", ct);
    *ld = ct;
    if (L.S == 0) return 0;
    Scratch sc;
    if ((e = pass_scratch(p, L, sc))) return e;
    switch (ct) {
        case 1: e = rhs_cols<1>(p, L, sc, x, x32, nx, b, s); break;
        case 2: e = rhs_cols<2>(p, L, sc, x, x32, nx, b, s); break;
        case 3: e = rhs_cols<3>(p, L, sc, x, x32, nx, b, s); break;
        case 4: e = rhs_cols<4>(p, L, sc, x, x32, nx, b, s); break;
        case 6: e = rhs_cols<6>(p, L, sc, x, x32, nx, b, s); break;
        case 8: e = rhs_cols<8>(p, L, sc, x, x32, nx, b, s); break;
        case 9: e = rhs_cols<9>(p, L, sc, x, x32, nx, b, s); break;
        case 12: e = rhs_cols<12>(p, L, sc, x, x32, nx, b, s); break;
        default: e = rhs_cols<16>(p, L, sc, x, x32, nx, b, s); break;
    }
    if (e) return e;
    FEHIP_CHECK(hipStreamSynchronize(s), "shared right-hand sides");  // the scratch is released on return
    return 0;
}

// column-tile width cap: widest of {16, 12, 9, 8, 6, 4, 3, 2, 1} whose small-dim LDS tables fit next to the
// row tile (<= 28 KB of staged + accumulated tables); 0 if none does
static int tile_cap(const Layout& L) {
    static const int widths[] = {16, 12, 9, 8, 6, 4, 3, 2, 1};
    for (int wdt : widths)
        if ((size_t)L.Gs * wdt * 16 <= 28 * 1024) return wdt;
    return 0;
}

int within_tile_cap(Panel* p, unsigned fe_mask) {
    unsigned mask = fe_mask ? fe_mask : ((1u << p->m) - 1u);
    mask &= (1u << p->m) - 1u;
    Layout L;
    if (make_layout(p, mask, L)) return 0;
    return tile_cap(L);
}

int within_transform(Panel* p, const WithinCols& wc, int ncols, const uint32_t* wm, int nslots,
                     const fehip_demean_opts& o, int to_work, WithinStats& ws, hipStream_t s) {
    unsigned mask = o.fe_mask ? o.fe_mask : ((1u << p->m) - 1u);
    mask &= (1u << p->m) - 1u;
    Layout L;
    int e = make_layout(p, mask, L);
    if (e) return e;
    ws = WithinStats(ncols);
    if (o.method != 0) return set_error("within transform: method %d is not available (0: PCG in FE space)", o.method);
    if (ncols <= 0) return 0;
    if (!wm) nslots = 1;
    if (nslots < 1 || nslots > 32) return set_error("1..32 request slots per solve (%d)", nslots);
    for (int c = 0; c < ncols; ++c)
        if (wc.slot[c] < 0 || wc.slot[c] >= nslots) return set_error("column %d: slot out of range", c);
    Scratch sc;
    const int ct_max = tile_cap(L);
    if (ct_max == 0) return set_error("small FE dims exceed the LDS budget (%d groups)", L.Gs);
    const int64_t S = L.S;
    const int64_t n = p->n;
    if ((e = sc.alloc(&sc.tab, (size_t)(S * kMaxCT * 7 + 16)))) return e;
    if ((e = sc.alloc(&sc.dg, (size_t)(S * 2 * nslots + 2)))) return e;
    for (int i = 0; i < L.nd; ++i)
        if (!L.sd[i].small) {
            const int64_t R = p->d[L.dims[i]].n_runs;
            if ((e = sc.alloc(&sc.runbuf[L.sd[i].large_slot], (size_t)(R * kMaxCT + 1)))) return e;
        }
    // the FP32 operator runs on panels whose primary segments all lie in wave tiles (the lean kernel's
    // domain); elsewhere a mixed-precision solve uses the FP64 operator
    if (o.precision == 1 && S > 0 && p->n_mtiles == 0 && p->n_long == 0) {
        for (int i = 0; i < L.nd; ++i)
            if (!L.sd[i].small) {
                const int64_t R = p->d[L.dims[i]].n_runs;
                if ((e = sc.alloc(&sc.runbuf32[L.sd[i].large_slot], (size_t)(R * kMaxCT + 4)))) return e;
            }
        if ((e = sc.alloc(&sc.p32, (size_t)(S * kMaxCT + 4)))) return e;
        if ((e = sc.alloc(&sc.r32, (size_t)(S * kMaxCT + 4)))) return e;
        if ((e = sc.alloc(&sc.bvec, (size_t)(S * kMaxCT + 4)))) return e;
        // padding entries of the FP32 rows are loaded with the row (never used): keep them finite
        FEHIP_CHECK(hipMemsetAsync(sc.p32, 0, sizeof(float) * (size_t)(S * kMaxCT + 4), s), "p32 clear");
    }
    const unsigned grid = apply_grid(p);
    if ((e = sc.alloc(&sc.small_part, (size_t)grid * (L.Gs > 0 ? L.Gs : 1) * kMaxCT))) return e;
    if ((e = sc.alloc(&sc.dump, (size_t)wave_grid(p) * (kBlock / kWave) * kMaxCT + kMaxCT))) return e;
    if ((e = sc.alloc(&sc.part, (size_t)kCgBlocks * kMaxCT * 2))) return e;
    if (wc.gram) {
        if (wc.gram_ld < ncols) return set_error("within transform: cross-product leading dim %lld < %d columns",
                                                  (long long)wc.gram_ld, ncols);
        const size_t nrec = grid + 1;
        if ((e = sc.alloc(&sc.gram_rec, (nrec + (nrec + kGramRun - 1) / kGramRun) * 256))) return e;
    }
    if ((e = sc.alloc(&sc.st, 1))) return e;
    sc.st_host = (CgState*)pinned_scratch(sizeof(CgState));
    if (!sc.st_host) return set_error("pinned host allocation failed");
    // Schur diagonal of every slot in use (an unweighted solve's diagonal depends only on the panel and
    // the absorbed dims, and is kept with the panel)
    if (S > 0 && !wm) {
        std::lock_guard<std::recursive_mutex> guard(p->mu);
        auto it = p->diag_cache.find(mask);
        if (it != p->diag_cache.end()) {
            sc.dg = it->second;
        } else {
            if ((e = slot_diagonal(p, L, sc, nullptr, ~0u, sc.dg, s))) return e;
            double* keep = nullptr;
            if ((e = dalloc(&keep, (size_t)S * 2, &p->bytes))) return e;
            FEHIP_CHECK(hipMemcpyAsync(keep, sc.dg, sizeof(double) * (size_t)S * 2, hipMemcpyDeviceToDevice, s),
                        "diagonal copy");
            p->diag_cache[mask] = keep;
            sc.dg = keep;
        }
    } else if (S > 0) {
        std::vector<int> used(nslots, 0);
        for (int c = 0; c < ncols; ++c) used[wc.slot[c]] = 1;
        for (int b = 0; b < nslots; ++b)
            if (used[b] && (e = slot_diagonal(p, L, sc, wm, 1u << b, sc.dg + (int64_t)b * S * 2, s))) return e;
    }
    // the run-level operator for the CG products where it applies (the tuner below times it)
    const int op_kind = table_operator_kind(p, L, wm, nslots);
    if (op_kind >= 0 && (e = run_couplings(p, L, mask, op_kind, sc, wm, s))) return e;
    // column tiles: the tuner's measured cost per width picks the split of ncols columns; with cross
    // products asked and at most ct_max columns, one tile holds them all
    std::vector<int> tiles;
    if (wc.gram && ncols <= ct_max) {
        for (int wdt : kWidths)
            if (wdt >= ncols) {
                tiles.push_back(wdt);
                break;
            }
    } else if ((e = plan_tiles(p, L, sc, mask, ncols, ct_max, tiles, s))) {
        return e;
    }
    int c0 = 0;
    int tile_id = 0;
    for (int ct : tiles) {
        const int nc = std::min(ct, ncols - c0);
        TileStats ts;
        switch (ct) {
            case 1: e = solve_tile<1>(p, L, sc, wc, c0, nc, wm, o, to_work, ts, s); break;
            case 2: e = solve_tile<2>(p, L, sc, wc, c0, nc, wm, o, to_work, ts, s); break;
            case 3: e = solve_tile<3>(p, L, sc, wc, c0, nc, wm, o, to_work, ts, s); break;
            case 4: e = solve_tile<4>(p, L, sc, wc, c0, nc, wm, o, to_work, ts, s); break;
            case 6: e = solve_tile<6>(p, L, sc, wc, c0, nc, wm, o, to_work, ts, s); break;
            case 8: e = solve_tile<8>(p, L, sc, wc, c0, nc, wm, o, to_work, ts, s); break;
            case 9: e = solve_tile<9>(p, L, sc, wc, c0, nc, wm, o, to_work, ts, s); break;
            case 12: e = solve_tile<12>(p, L, sc, wc, c0, nc, wm, o, to_work, ts, s); break;
            default: e = solve_tile<16>(p, L, sc, wc, c0, nc, wm, o, to_work, ts, s); break;
        }
        if (e) return e;
        ws.iterations = std::max(ws.iterations, ts.iterations);
        ws.converged = ws.converged && ts.converged;
        ws.max_update = std::max(ws.max_update, ts.max_update);
        ws.seconds += ts.seconds;
        ws.apply_launches += ts.apply_launches;
        for (int c = 0; c < nc; ++c) {
            ws.col_iters[c0 + c] = ts.col_iters[c];
            ws.col_converged[c0 + c] = ts.col_converged[c];
            ws.col_max_update[c0 + c] = ts.col_max_update[c];
            ws.col_tile[c0 + c] = tile_id;
            ws.col_refines[c0 + c] = ts.refine_launches;
        }
        ++tile_id;
        const int64_t id_bytes = n * 4 * L.nd + (int64_t)(p->d[p->P].G + 1) * 8 + (wm ? 4 * n : 0);
        ws.bytes += ts.apply_launches * id_bytes + (int64_t)3 * n * 8 * nc;
        c0 += nc;
    }
    return 0;
}

}  // namespace fehip

using namespace fehip;

extern "C" int fehip_demean(fehip_panel* panel, double* d_cols, int n_cols, int64_t ld, const uint8_t* d_weights,
                            const fehip_demean_opts* opts, fehip_demean_stats* stats) {
    (void)hipGetLastError();  // start from a clean error state
    if (!panel || !opts) return set_error("fehip_demean: null argument");
    Panel* p = reinterpret_cast<Panel*>(panel);
    if (n_cols < 0 || (n_cols > 0 && (!d_cols || ld < p->n))) return set_error("fehip_demean: bad columns");
    WithinCols wc(n_cols);
    for (int c = 0; c < n_cols; ++c) {
        wc.x[c] = d_cols + (int64_t)c * ld;
        wc.out[c] = d_cols + (int64_t)c * ld;
        wc.slot[c] = 0;
    }
    uint32_t* wm = nullptr;
    if (d_weights) {
        int e = talloc(&wm, (size_t)p->n);
        if (e) return e;
        e = build_mask_words(p, &d_weights, 1, wm, 0);
        if (e) {
            cache_free(wm);
            return e;
        }
    }
    WithinStats ws(n_cols);
    int e = within_transform(p, wc, n_cols, wm, 1, *opts, 0, ws, 0);
    cache_free(wm);
    if (e) return e;
    if (stats) {
        stats->iterations = ws.iterations;
        stats->converged = ws.converged;
        stats->max_update = ws.max_update;
        stats->seconds = ws.seconds;
        stats->bytes_moved = ws.bytes;
    }
    return set_ok();
}
