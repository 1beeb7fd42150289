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
// Stopping rule: max over groups of |r_g / n_g|, the size of the group-mean update the next
// alternating projection would make -- the reference's criterion (fe_gpu_demean.f90:196-211,294).
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "kstats.hpp"
#include "panel.hpp"
#include "within.hpp"

namespace fehip {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxScatter = 6;
constexpr int kMaxLarge = 3;

enum Mode { kIter = 0, kData = 1, kDiag = 2, kFinal = 3 };

struct ScatterDim {
    const int* ids;
    int G;
    int small;
    int64_t tab_off;             // group offset in table space
    int lds_off;                 // small: group offset in LDS
    int large_slot;              // large: 0..kMaxLarge-1
    const int64_t* seg_run_off;  // large
    double* runbuf;              // large: [n_runs][CT]
};

struct ApplyArgs {
    const int64_t* seg_off;
    int64_t GP;
    int elim;
    const uint8_t* w;       // panel-order weights or null
    int nd;
    ScatterDim sd[kMaxScatter];
    int Gs;                 // small groups in LDS
    int diag_dim;           // kDiag: the scatter dim processed
    const double* p;        // kIter: [S][CT]
    const int* active;      // kIter: per-column CG activity (nullptr = all active)
    const double* xc[16];   // kData: caller columns of this tile (caller row order)
    double* oc[16];         // finalize: output columns
    int ncols;              // valid columns in this tile
    const int* perm;
    double* small_part;     // [blocks][Gs][CT]
    const int64_t* tile_seg;   // [2 * n_tiles] segment range of each tile
    int64_t n_tiles;
    const int64_t* long_segs;  // segments longer than a tile
    int64_t n_long;
    int part_base;             // first partial block written by this launch
    const double* alpha;       // finalize: solved tables [S][CT]
    int to_work;               // finalize: 1 = panel-order work columns, 0 = caller columns in place
};

template <int CT, int MODE>
__device__ __forceinline__ void row_values(const ApplyArgs& a, int64_t r, const double* sp, double (&v)[CT]) {
    if constexpr (MODE == kData) {
        const int64_t src = a.perm ? a.perm[r] : r;
#pragma unroll
        for (int c = 0; c < CT; ++c) v[c] = c < a.ncols ? a.xc[c][src] : 0.0;
    } else {
#pragma unroll
        for (int c = 0; c < CT; ++c) v[c] = 0.0;
        for (int o = 0; o < a.nd; ++o) {
            const ScatterDim& d = a.sd[o];
            const int g = d.ids[r];
            if (d.small) {
                const double* t = sp + (int64_t)(d.lds_off + g) * CT;
#pragma unroll
                for (int c = 0; c < CT; ++c) v[c] += t[c];
            } else {
                const double* t = a.p + (d.tab_off + g) * CT;
#pragma unroll
                for (int c = 0; c < CT; ++c) v[c] += t[c];
            }
        }
    }
}

// One thread per primary segment; see the file comment for the three modes.
template <int CT, int MODE>
__global__ __launch_bounds__(kBlock) void apply_kernel(ApplyArgs a) {
    extern __shared__ double lds[];
    double* sacc = lds;                                   // [Gs][CT]
    double* sp = lds + (size_t)a.Gs * CT;                  // [Gs][CT] (kIter)
    for (int e = threadIdx.x; e < a.Gs * CT; e += blockDim.x) sacc[e] = 0.0;
    if constexpr (MODE == kIter) {
        for (int o = 0; o < a.nd; ++o) {
            const ScatterDim& d = a.sd[o];
            if (!d.small) continue;
            for (int e = threadIdx.x; e < d.G * CT; e += blockDim.x)
                sp[(size_t)d.lds_off * CT + e] = a.p[d.tab_off * CT + e];
        }
    }
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; li < a.n_long; li += stride) {
        const int64_t s = a.long_segs[li];
        const int64_t r0 = a.seg_off[s], r1 = a.seg_off[s + 1];
        double mean[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) mean[c] = 0.0;
        double nw = 0.0;
        if (a.elim || MODE == kDiag) {
            double acc[CT];
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[c] = 0.0;
            for (int64_t r = r0; r < r1; ++r) {
                if (a.w && !a.w[r]) continue;
                nw += 1.0;
                if constexpr (MODE != kDiag) {
                    double v[CT];
                    row_values<CT, MODE>(a, r, sp, v);
#pragma unroll
                    for (int c = 0; c < CT; ++c) acc[c] += v[c];
                }
            }
            if (nw > 0.0)
#pragma unroll
                for (int c = 0; c < CT; ++c) mean[c] = acc[c] / nw;
        }
        int cur[kMaxLarge];
        int64_t rk[kMaxLarge];
        double racc[kMaxLarge][CT];
#pragma unroll
        for (int j = 0; j < kMaxLarge; ++j) {
            cur[j] = -1;
            rk[j] = 0;
#pragma unroll
            for (int c = 0; c < CT; ++c) racc[j][c] = 0.0;
        }
        for (int o = 0; o < a.nd; ++o)
            if (!a.sd[o].small) rk[a.sd[o].large_slot] = a.sd[o].seg_run_off[s];
        for (int64_t r = r0; r < r1; ++r) {
            const bool on = !a.w || a.w[r];
            double u[CT];
            if constexpr (MODE == kDiag) {
                // u0: Schur-diagonal contribution of this row to its group of dim diag_dim,
                // u1: its weight.  c_r = weight of the segment's rows sharing the group.
                double cr = 0.0;
                if (on && a.elim) {
                    const int* ids = a.sd[a.diag_dim].ids;
                    const int g = ids[r];
                    for (int64_t q = r0; q < r1; ++q) cr += (ids[q] == g && (!a.w || a.w[q])) ? 1.0 : 0.0;
                }
#pragma unroll
                for (int c = 0; c < CT; ++c) u[c] = 0.0;
                if (on) {
                    u[0] = a.elim ? 1.0 - cr / nw : 1.0;
                    if (CT > 1) u[1 % CT] = 1.0;
                }
            } else {
                double v[CT];
                row_values<CT, MODE>(a, r, sp, v);
#pragma unroll
                for (int c = 0; c < CT; ++c) u[c] = on ? v[c] - mean[c] : 0.0;
            }
            for (int o = 0; o < a.nd; ++o) {
                if (MODE == kDiag && o != a.diag_dim) continue;
                const ScatterDim& d = a.sd[o];
                const int g = d.ids[r];
                if (d.small) {
                    if (on) {
                        double* t = sacc + (size_t)(d.lds_off + g) * CT;
#pragma unroll
                        for (int c = 0; c < CT; ++c) lds_add(&t[c], u[c]);
                    }
                } else {
                    const int j = d.large_slot;
                    if (g != cur[j]) {
                        if (cur[j] >= 0) {
                            double* dst = d.runbuf + rk[j] * CT;
#pragma unroll
                            for (int c = 0; c < CT; ++c) dst[c] = racc[j][c];
                            ++rk[j];
                        }
                        cur[j] = g;
#pragma unroll
                        for (int c = 0; c < CT; ++c) racc[j][c] = 0.0;
                    }
#pragma unroll
                    for (int c = 0; c < CT; ++c) racc[j][c] += u[c];
                }
            }
        }
        for (int o = 0; o < a.nd; ++o) {
            if (MODE == kDiag && o != a.diag_dim) continue;
            const ScatterDim& d = a.sd[o];
            if (d.small) continue;
            const int j = d.large_slot;
            if (cur[j] >= 0) {
                double* dst = d.runbuf + rk[j] * CT;
#pragma unroll
                for (int c = 0; c < CT; ++c) dst[c] = racc[j][c];
            }
        }
    }
    __syncthreads();
    double* part = a.small_part + (size_t)(a.part_base + blockIdx.x) * a.Gs * CT;
    for (int e = threadIdx.x; e < a.Gs * CT; e += blockDim.x) part[e] = sacc[e];
}


// Row-tiled kernel: one workgroup per tile of whole primary segments (<= kTileRows rows), one row
// per thread.  Phase 1 loads each row's values (coalesced) into LDS, phase 2 removes the segment
// means (one thread per segment, LDS only), phase 3 scatters the rows into the other dims' tables
// (LDS atomics for small dims; per-run partials for large dims, runs ranked with a block scan) or,
// for kFinal, writes the transformed rows out.
template <int CT, int MODE>
__global__ __launch_bounds__(kBlock) void tile_kernel(ApplyArgs a) {
    __shared__ double sv[kTileRows * CT];
    __shared__ int gd[kTileRows];            // kDiag: group of the processed dim
    __shared__ unsigned char wr[kTileRows];  // row weight (0/1), 0 past the tile end
    __shared__ unsigned char sb[kTileRows];  // row starts a segment
    __shared__ unsigned char hf[kTileRows];  // row starts a run (large-dim scatter)
    __shared__ short rs[kTileRows], re[kTileRows];  // kDiag: segment range of each row
    __shared__ int wcnt[kBlock / kWave];
    extern __shared__ double dyn[];
    double* sacc = dyn;                          // [Gs][CT]
    double* sp = dyn + (size_t)a.Gs * CT;        // [Gs][CT] staged p (kIter) or alpha (kFinal)
    const int j = threadIdx.x;
    const int lane = j & (kWave - 1), wv = j / kWave;
    for (int e = j; e < a.Gs * CT; e += blockDim.x) sacc[e] = 0.0;
    if constexpr (MODE == kIter || MODE == kFinal) {
        const double* tab = MODE == kIter ? a.p : a.alpha;
        for (int o = 0; o < a.nd; ++o) {
            const ScatterDim& d = a.sd[o];
            if (!d.small) continue;
            for (int e = j; e < d.G * CT; e += blockDim.x) sp[(size_t)d.lds_off * CT + e] = tab[d.tab_off * CT + e];
        }
    }
    __syncthreads();
    for (int64_t t = blockIdx.x; t < a.n_tiles; t += gridDim.x) {
        const int64_t s0 = a.tile_seg[2 * t], s1 = a.tile_seg[2 * t + 1];
        const int64_t R0 = a.seg_off[s0];
        const int nr = (int)(a.seg_off[s1] - R0);
        const int64_t r = R0 + j;
        const bool valid = j < nr;
        const unsigned char w = valid ? (a.w ? a.w[r] : 1) : 0;
        wr[j] = w;
        sb[j] = 0;
        // ---- phase 1: row values
        if constexpr (MODE == kDiag) {
            gd[j] = valid ? a.sd[a.diag_dim].ids[r] : -1;
        } else {
            double v[CT];
#pragma unroll
            for (int c = 0; c < CT; ++c) v[c] = 0.0;
            if (valid) {
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
                            for (int c = 0; c < CT; ++c) v[c] += sign * sp[base + c];
                        } else {
                            const int64_t base = (d.tab_off + g) * CT;
#pragma unroll
                            for (int c = 0; c < CT; ++c) v[c] += sign * tab[base + c];
                        }
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < CT; ++c) sv[j * CT + c] = w ? v[c] : 0.0;
        }
        __syncthreads();
        // ---- phase 2: segment means (one thread per segment)
        for (int64_t k = j; k < s1 - s0; k += blockDim.x) {
            const int b = (int)(a.seg_off[s0 + k] - R0), e = (int)(a.seg_off[s0 + k + 1] - R0);
            if (b < e) sb[b] = 1;
            if constexpr (MODE == kDiag) {
                for (int q = b; q < e; ++q) {
                    rs[q] = (short)b;
                    re[q] = (short)e;
                }
            } else {
                if (a.elim && b < e) {
                    double acc[CT];
#pragma unroll
                    for (int c = 0; c < CT; ++c) acc[c] = 0.0;
                    int nw = 0;
                    for (int q = b; q < e; ++q) {
                        nw += wr[q];
#pragma unroll
                        for (int c = 0; c < CT; ++c) acc[c] += sv[q * CT + c];
                    }
                    if (nw > 0) {
                        const double inv = 1.0 / (double)nw;
                        for (int q = b; q < e; ++q)
                            if (wr[q])
#pragma unroll
                                for (int c = 0; c < CT; ++c) sv[q * CT + c] -= acc[c] * inv;
                    }
                }
            }
        }
        __syncthreads();
        if constexpr (MODE == kDiag) {
            // u0: this row's contribution to the Schur diagonal of its group, u1: its weight
            double u0 = 0.0, u1 = 0.0;
            if (valid && w) {
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
            if (CT > 1) sv[j * CT + (1 % CT)] = u1;
            __syncthreads();
        }
        // ---- phase 3
        if constexpr (MODE == kFinal) {
            if (valid) {
                const int64_t src = a.perm ? a.perm[r] : r;
                for (int c = 0; c < a.ncols; ++c) {
                    if (a.to_work)
                        a.oc[c][r] = w ? sv[j * CT + c] : 0.0;
                    else if (w)
                        a.oc[c][src] = sv[j * CT + c];
                }
            }
        } else {
            for (int o = 0; o < a.nd; ++o) {
                if (MODE == kDiag && o != a.diag_dim) continue;
                const ScatterDim& d = a.sd[o];
                if (d.small) {
                    if (valid && w) {
                        const int g = MODE == kDiag ? gd[j] : d.ids[r];
                        double* tq = sacc + (size_t)(d.lds_off + g) * CT;
#pragma unroll
                        for (int c = 0; c < CT; ++c) lds_add(&tq[c], sv[j * CT + c]);
                    }
                    continue;
                }
                // runs: a new run starts at a segment start or where the id changes
                const int g = valid ? (MODE == kDiag ? gd[j] : d.ids[r]) : -1;
                const int gp = (valid && j > 0) ? (MODE == kDiag ? gd[j - 1] : d.ids[r - 1]) : -1;
                const bool head = valid && (j == 0 || sb[j] || g != gp);
                hf[j] = head;
                const unsigned long long bal = __ballot(head);
                if (lane == 0) wcnt[wv] = __popcll(bal);
                __syncthreads();
                if (head) {
                    int rank = __popcll(bal & ((1ull << lane) - 1ull));
                    for (int q = 0; q < wv; ++q) rank += wcnt[q];
                    double acc[CT];
#pragma unroll
                    for (int c = 0; c < CT; ++c) acc[c] = sv[j * CT + c];
                    for (int q = j + 1; q < nr && !hf[q]; ++q)
#pragma unroll
                        for (int c = 0; c < CT; ++c) acc[c] += sv[q * CT + c];
                    double* dst = d.runbuf + (d.seg_run_off[s0] + rank) * CT;
#pragma unroll
                    for (int c = 0; c < CT; ++c) dst[c] = acc[c];
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if constexpr (MODE != kFinal) {
        double* part = a.small_part + (size_t)(a.part_base + blockIdx.x) * a.Gs * CT;
        for (int e = j; e < a.Gs * CT; e += blockDim.x) part[e] = sacc[e];
    }
}


// Row tile (kIter, kData, kFinal): one workgroup per tile of whole primary segments, one row per
// thread.  Rows are staged in LDS; segment means are taken by one thread per (segment, column)
// and run sums by one thread per (run, column), so no thread walks a long serial chain and no
// cross-lane shuffles are needed.  Run partials of a tile are contiguous in the run buffer and
// leave in one coalesced store; small-dim tables take LDS atomics.
template <int CT, int MODE>
__global__ __launch_bounds__(kBlock) void tile_fast_kernel(ApplyArgs a) {
    if constexpr (MODE == kIter) {
        if (a.active) {  // every column converged: nothing to apply (the host polls periodically)
            bool any = false;
#pragma unroll
            for (int c = 0; c < CT; ++c) any |= a.active[c] != 0;
            if (!any) return;
        }
    }
    __shared__ double sv[kTileRows * CT];     // row values
    __shared__ unsigned char wr[kTileRows];   // row weight
    __shared__ int soff[kTileRows + 1];       // local row offsets of the tile's segments
    __shared__ int rstart[kTileRows + 1];     // first row of each run
    __shared__ int wcnt[kBlock / kWave];
    extern __shared__ double dyn[];
    double* sacc = dyn;                       // [Gs][CT]
    double* sp = dyn + (size_t)a.Gs * CT;     // [Gs][CT]
    const int j = threadIdx.x;
    const int lane = j & (kWave - 1), wv = j / kWave;
    for (int e = j; e < a.Gs * CT; e += blockDim.x) sacc[e] = 0.0;
    if constexpr (MODE == kIter || MODE == kFinal) {
        const double* tab = MODE == kIter ? a.p : a.alpha;
        for (int o = 0; o < a.nd; ++o) {
            const ScatterDim& d = a.sd[o];
            if (!d.small) continue;
            for (int e = j; e < d.G * CT; e += blockDim.x) sp[(size_t)d.lds_off * CT + e] = tab[d.tab_off * CT + e];
        }
    }
    for (int64_t t = blockIdx.x; t < a.n_tiles; t += gridDim.x) {
        const int64_t s0 = a.tile_seg[2 * t], s1 = a.tile_seg[2 * t + 1];
        const int nseg = (int)(s1 - s0);
        const int64_t R0 = a.seg_off[s0];
        __syncthreads();  // previous tile done with the LDS arrays (and the staged tables are ready)
        for (int k = j; k <= nseg; k += blockDim.x) soff[k] = (int)(a.seg_off[s0 + k] - R0);
        const int nr = (int)(a.seg_off[s1] - R0);
        const bool valid = j < nr;
        const int64_t r = R0 + j;
        const bool on = valid && (!a.w || a.w[r]);
        int gid[kMaxScatter];
        double v[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) v[c] = 0.0;
        for (int o = 0; o < a.nd; ++o) gid[o] = valid ? a.sd[o].ids[r] : -1;
        if (on) {
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
                    if (d.small) {
                        const size_t base = (size_t)(d.lds_off + gid[o]) * CT;
#pragma unroll
                        for (int c = 0; c < CT; ++c) v[c] += sign * sp[base + c];
                    } else {
                        const int64_t base = (d.tab_off + gid[o]) * CT;
#pragma unroll
                        for (int c = 0; c < CT; ++c) v[c] += sign * tab[base + c];
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) sv[j * CT + c] = v[c];
        wr[j] = on ? 1 : 0;
        __syncthreads();
        if (a.elim) {
            // one thread per (segment, column): subtract the weighted segment mean
            for (int e = j; e < nseg * CT; e += blockDim.x) {
                const int k = e / CT, c = e - k * CT;
                const int b = soff[k], f = soff[k + 1];
                double sm = 0.0;
                int nw = 0;
                for (int q = b; q < f; ++q) {
                    sm += sv[q * CT + c];
                    nw += wr[q];
                }
                if (nw > 0) {
                    const double mean = sm / (double)nw;
                    for (int q = b; q < f; ++q)
                        if (wr[q]) sv[q * CT + c] -= mean;
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
                    if (a.to_work)
                        a.oc[c][r] = sv[j * CT + c];
                    else if (on)
                        a.oc[c][src] = sv[j * CT + c];
                }
            }
        } else {
            for (int o = 0; o < a.nd; ++o) {
                const ScatterDim& d = a.sd[o];
                if (d.small) {
                    if (on) {
                        double* tq = sacc + (size_t)(d.lds_off + gid[o]) * CT;
#pragma unroll
                        for (int c = 0; c < CT; ++c) lds_add(&tq[c], sv[j * CT + c]);
                    }
                    continue;
                }
                // runs: a new run starts at a segment start or where the id changes
                int gp = __shfl_up(gid[o], 1, kWave);
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
                const bool rhead = valid && (j == 0 || segstart || gid[o] != gp);
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
                double* dst = d.runbuf + d.seg_run_off[s0] * CT;
                for (int e = j; e < total * CT; e += blockDim.x) {
                    const int k = e / CT, c = e - k * CT;
                    double sm = 0.0;
                    for (int q = rstart[k]; q < rstart[k + 1]; ++q) sm += sv[q * CT + c];
                    dst[e] = sm;
                }
                __syncthreads();
            }
        }
    }
    if constexpr (MODE != kFinal) {
        __syncthreads();
        double* part = a.small_part + (size_t)(a.part_base + blockIdx.x) * a.Gs * CT;
        for (int e = j; e < a.Gs * CT; e += blockDim.x) part[e] = sacc[e];
    }
}


// q[tab_off + g][c] = sum of the run partials of group g (runs in group order)
template <int CT>
__global__ void reduce_runs_kernel(const double* __restrict__ runbuf, const int* __restrict__ order,
                                   const int64_t* __restrict__ goff, int G, double* q) {
    const int64_t total = (int64_t)G * CT;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int g = (int)(e / CT), c = (int)(e % CT);
        double s = 0.0;
        for (int64_t k = goff[g]; k < goff[g + 1]; ++k) s += runbuf[(int64_t)order[k] * CT + c];
        q[e] = s;
    }
}

// q[tab_off + g][c] = sum over blocks of the LDS partials
template <int CT>
__global__ __launch_bounds__(kBlock) void reduce_small_kernel(const double* __restrict__ part, int nblk, int Gs,
                                                               int lds_off, int G, double* q) {
    // one workgroup per table entry: the threads stride over the per-block partials in a fixed
    // order and combine them with a fixed-shape tree
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

// ---- CG in table space ([S][CT], one CG per column) ----
struct CgState {
    double rz[16], alpha[16], beta[16], maxn[16];
    int active[16], iters[16];
};

template <int CT>
__global__ __launch_bounds__(kBlock) void cg_init_kernel(int64_t S, const double* __restrict__ r,
                                                         const double* __restrict__ minv,
                                                         const double* __restrict__ wcnt, double* z, double* p,
                                                         double* x, double* part_rz, double* part_max) {
    __shared__ double ws[2][kBlock / kWave][CT];
    double rz[CT], mx[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) rz[c] = mx[c] = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < S; g += (int64_t)gridDim.x * blockDim.x) {
        const double mi = minv[g], n = wcnt[g];
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const double rv = r[g * CT + c];
            const double zv = mi * rv;
            z[g * CT + c] = zv;
            p[g * CT + c] = zv;
            x[g * CT + c] = 0.0;
            rz[c] += rv * zv;
            if (n > 0.0) mx[c] = fmax(mx[c], fabs(rv) / n);
        }
    }
    const int wv = threadIdx.x / kWave, ln = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        double a = wave_sum(rz[c]), b = wave_max(mx[c]);
        if (ln == 0) {
            ws[0][wv][c] = a;
            ws[1][wv][c] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < CT) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) {
            a += ws[0][w][threadIdx.x];
            b = fmax(b, ws[1][w][threadIdx.x]);
        }
        part_rz[blockIdx.x * CT + threadIdx.x] = a;
        part_max[blockIdx.x * CT + threadIdx.x] = b;
    }
}

template <int CT>
__global__ void cg_start_scalars(const double* part_rz, const double* part_max, int nblk, double tol, int ncols,
                                 CgState* st) {
    const int c = threadIdx.x;
    if (c >= CT) return;
    double rz = 0.0, mx = 0.0;
    for (int b = 0; b < nblk; ++b) {
        rz += part_rz[b * CT + c];
        mx = fmax(mx, part_max[b * CT + c]);
    }
    st->rz[c] = rz;
    st->maxn[c] = mx;
    st->iters[c] = 0;
    st->active[c] = (c < ncols) && (mx >= tol) && rz > 0.0;
    st->alpha[c] = st->beta[c] = 0.0;
}

template <int CT>
__global__ __launch_bounds__(kBlock) void cg_pq_kernel(int64_t S, const double* __restrict__ p,
                                                       const double* __restrict__ q, double* part) {
    __shared__ double ws[kBlock / kWave][CT];
    double a[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) a[c] = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < S; g += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < CT; ++c) a[c] += p[g * CT + c] * q[g * CT + c];
    const int wv = threadIdx.x / kWave, ln = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        double t = wave_sum(a[c]);
        if (ln == 0) ws[wv][c] = t;
    }
    __syncthreads();
    if (threadIdx.x < CT) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += ws[w][threadIdx.x];
        part[blockIdx.x * CT + threadIdx.x] = t;
    }
}

template <int CT>
__global__ void cg_alpha_scalars(const double* part, int nblk, CgState* st) {
    const int c = threadIdx.x;
    if (c >= CT) return;
    double pq = 0.0;
    for (int b = 0; b < nblk; ++b) pq += part[b * CT + c];
    st->alpha[c] = (st->active[c] && pq > 0.0) ? st->rz[c] / pq : 0.0;
}

template <int CT>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(int64_t S, const CgState* __restrict__ st,
                                                           const double* __restrict__ p, const double* __restrict__ q,
                                                           const double* __restrict__ minv,
                                                           const double* __restrict__ wcnt, double* x, double* r,
                                                           double* z, double* part_rz, double* part_max) {
    __shared__ double ws[2][kBlock / kWave][CT];
    double al[CT], rz[CT], mx[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        al[c] = st->alpha[c];
        rz[c] = mx[c] = 0.0;
    }
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < S; g += (int64_t)gridDim.x * blockDim.x) {
        const double mi = minv[g], n = wcnt[g];
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const int64_t e = g * CT + c;
            double rv = r[e];
            if (al[c] != 0.0) {
                x[e] += al[c] * p[e];
                rv -= al[c] * q[e];
                r[e] = rv;
            }
            const double zv = mi * rv;
            z[e] = zv;
            rz[c] += rv * zv;
            if (n > 0.0) mx[c] = fmax(mx[c], fabs(rv) / n);
        }
    }
    const int wv = threadIdx.x / kWave, ln = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        double a = wave_sum(rz[c]), b = wave_max(mx[c]);
        if (ln == 0) {
            ws[0][wv][c] = a;
            ws[1][wv][c] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < CT) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) {
            a += ws[0][w][threadIdx.x];
            b = fmax(b, ws[1][w][threadIdx.x]);
        }
        part_rz[blockIdx.x * CT + threadIdx.x] = a;
        part_max[blockIdx.x * CT + threadIdx.x] = b;
    }
}

template <int CT>
__global__ void cg_beta_scalars(const double* part_rz, const double* part_max, int nblk, double tol, CgState* st) {
    const int c = threadIdx.x;
    if (c >= CT) return;
    if (!st->active[c]) {
        st->beta[c] = 0.0;
        return;
    }
    double rz = 0.0, mx = 0.0;
    for (int b = 0; b < nblk; ++b) {
        rz += part_rz[b * CT + c];
        mx = fmax(mx, part_max[b * CT + c]);
    }
    st->iters[c] += 1;
    st->maxn[c] = mx;
    st->beta[c] = st->rz[c] > 0.0 ? rz / st->rz[c] : 0.0;
    st->rz[c] = rz;
    if (mx < tol || !(rz > 0.0)) st->active[c] = 0;
}

template <int CT>
__global__ void cg_pdir_kernel(int64_t S, const CgState* __restrict__ st, const double* __restrict__ z, double* p) {
    double be[CT];
    int act[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        be[c] = st->beta[c];
        act[c] = st->active[c];
    }
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < S; g += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < CT; ++c)
            if (act[c]) p[g * CT + c] = z[g * CT + c] + be[c] * p[g * CT + c];
}

__global__ void make_minv_kernel(int64_t S, const double* __restrict__ dg, double* minv, double* wcnt_out,
                                 const double* __restrict__ wc) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < S; g += (int64_t)gridDim.x * blockDim.x) {
        const double d = dg[g * 2], n = wc[g * 2 + 1];
        minv[g] = (n > 0.0 && d > 1e-9 * n) ? 1.0 / d : 0.0;
        wcnt_out[g] = n;
    }
}

// sa = sum_o alpha_o[id_o(r)] (small dims from the LDS copy, large dims from global memory)
template <int CT>
__device__ __forceinline__ void alpha_sum(const ApplyArgs& a, int64_t r, const double* sp, const double* alpha,
                                          double (&sa)[CT]) {
#pragma unroll
    for (int c = 0; c < CT; ++c) sa[c] = 0.0;
    for (int o = 0; o < a.nd; ++o) {
        const ScatterDim& d = a.sd[o];
        const int g = d.ids[r];
        if (d.small) {
            const size_t base = (size_t)(d.lds_off + g) * CT;
#pragma unroll
            for (int c = 0; c < CT; ++c) sa[c] += sp[base + c];
        } else {
            const int64_t base = (d.tab_off + g) * CT;
#pragma unroll
            for (int c = 0; c < CT; ++c) sa[c] += alpha[base + c];
        }
    }
}

// out = M_P (x - D_o a) on weighted rows; to_work: write panel order into work (zeros for
// unweighted rows), else in place into the caller's columns.
template <int CT>
__global__ __launch_bounds__(kBlock) void finalize_kernel(ApplyArgs a, const double* __restrict__ alpha,
                                                          int to_work) {
    extern __shared__ double sp[];
    for (int o = 0; o < a.nd; ++o) {
        const ScatterDim& d = a.sd[o];
        if (!d.small) continue;
        for (int e = threadIdx.x; e < d.G * CT; e += blockDim.x) sp[(size_t)d.lds_off * CT + e] = alpha[d.tab_off * CT + e];
    }
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; li < a.n_long; li += stride) {
        const int64_t s = a.long_segs[li];
        const int64_t r0 = a.seg_off[s], r1 = a.seg_off[s + 1];
        double mean[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) mean[c] = 0.0;
        if (a.elim) {
            double acc[CT];
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[c] = 0.0;
            double nw = 0.0;
            for (int64_t r = r0; r < r1; ++r) {
                if (a.w && !a.w[r]) continue;
                nw += 1.0;
                const int64_t src = a.perm ? a.perm[r] : r;
                double sa[CT];
                alpha_sum<CT>(a, r, sp, alpha, sa);
#pragma unroll
                for (int c = 0; c < CT; ++c) acc[c] += (c < a.ncols ? a.xc[c][src] : 0.0) - sa[c];
            }
            if (nw > 0.0)
#pragma unroll
                for (int c = 0; c < CT; ++c) mean[c] = acc[c] / nw;
        }
        for (int64_t r = r0; r < r1; ++r) {
            const bool on = !a.w || a.w[r];
            const int64_t src = a.perm ? a.perm[r] : r;
            if (!on) {
                if (to_work)
                    for (int c = 0; c < a.ncols; ++c) a.oc[c][r] = 0.0;
                continue;
            }
            double sa[CT];
            alpha_sum<CT>(a, r, sp, alpha, sa);
#pragma unroll
            for (int c = 0; c < CT; ++c)
                if (c < a.ncols) a.oc[c][to_work ? r : src] = a.xc[c][src] - sa[c] - mean[c];
        }
    }
}

__global__ void weights_to_panel_kernel(const uint8_t* __restrict__ w, const int* __restrict__ perm, int64_t n,
                                        uint8_t* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = w[perm ? perm[i] : i] ? 1 : 0;
}

// ------------------------------------------------------------------------------------------
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
        L.S += s.G;
        if (s.small) {
            s.lds_off = L.Gs;
            L.Gs += s.G;
            s.large_slot = -1;
        } else {
            if (large == kMaxLarge) return set_error("more than %d large non-primary FE dims", kMaxLarge);
            s.large_slot = large++;
            s.seg_run_off = p->d[d].seg_run_off;
            s.lds_off = 0;
        }
        L.dims[L.nd++] = d;
    }
    return 0;
}

template <int CT>
size_t apply_lds(const Layout& L, int mode) {
    return sizeof(double) * (size_t)L.Gs * CT * (mode == kIter ? 2 : 1);
}

// dynamic LDS of the tile kernel: small-dim accumulators and staged tables
template <int CT>
size_t tile_lds(const Layout& L) {
    return sizeof(double) * (size_t)L.Gs * CT * 2;
}

unsigned tile_grid(Panel* p) {
    const int64_t cap = (int64_t)num_cus() * 8;
    const int64_t g = p->n_tiles < cap ? p->n_tiles : cap;
    return (unsigned)(g < 1 ? 1 : g);
}

unsigned long_grid(Panel* p) { return p->n_long > 0 ? stream_grid((size_t)p->n_long, kBlock, 2) : 0; }

unsigned apply_grid(Panel* p) { return tile_grid(p) + long_grid(p); }


// Launch one apply (mode) and reduce it into q (table space [S][CT]).
template <int CT>
int apply_and_reduce(Panel* p, const Layout& L, ApplyArgs& a, int mode, double* q, hipStream_t s) {
    const unsigned gt = tile_grid(p), gl = long_grid(p);
    const unsigned grid = gt + gl;
    const int sid = mode == kIter ? kStatApplyIter : kStatApplyData;
    const bool timed = mode != kDiag && kstats_enabled();
    if (timed) kstat_begin(sid, s);
    a.part_base = 0;
    const size_t lt = tile_lds<CT>(L);
    if (mode == kIter)
        hipLaunchKernelGGL((tile_fast_kernel<CT, kIter>), dim3(gt), dim3(kBlock), lt, s, a);
    else if (mode == kData)
        hipLaunchKernelGGL((tile_fast_kernel<CT, kData>), dim3(gt), dim3(kBlock), lt, s, a);
    else
        hipLaunchKernelGGL((tile_kernel<CT, kDiag>), dim3(gt), dim3(kBlock), lt, s, a);
    FEHIP_LAUNCH_CHECK("tile kernel");
    if (gl > 0) {
        ApplyArgs b = a;
        b.part_base = (int)gt;
        const size_t lds = apply_lds<CT>(L, mode);
        if (mode == kIter)
            hipLaunchKernelGGL((apply_kernel<CT, kIter>), dim3(gl), dim3(kBlock), lds, s, b);
        else if (mode == kData)
            hipLaunchKernelGGL((apply_kernel<CT, kData>), dim3(gl), dim3(kBlock), lds, s, b);
        else
            hipLaunchKernelGGL((apply_kernel<CT, kDiag>), dim3(gl), dim3(kBlock), lds, s, b);
        FEHIP_LAUNCH_CHECK("long-segment kernel");
    }
    if (timed) {
        // algorithmic bytes: id streams, segment offsets, weights, table reads, run-partial and
        // LDS-partial writes (+ the data columns and permutation for the right-hand side)
        const double n = (double)p->n;
        double bytes = n * 4.0 * L.nd + 8.0 * (double)(p->d[p->P].G + 1) + (a.w ? n : 0.0) +
                       8.0 * (double)grid * L.Gs * CT;
        for (int o = 0; o < L.nd; ++o) {
            const double G = (double)L.sd[o].G;
            bytes += 8.0 * CT * G;
            if (!L.sd[o].small) bytes += 8.0 * CT * (double)p->d[L.dims[o]].n_runs;
        }
        if (mode == kData) bytes += n * 8.0 * a.ncols + (p->perm ? 4.0 * n : 0.0);
        kstat_end(sid, s, bytes);
    }
    for (int o = 0; o < L.nd; ++o) {
        if (mode == kDiag && o != a.diag_dim) continue;
        const ScatterDim& d = L.sd[o];
        double* qo = q + d.tab_off * CT;
        if (d.small) {
            hipLaunchKernelGGL((reduce_small_kernel<CT>), dim3((unsigned)(d.G * CT)), dim3(kBlock), 0, s,
                               a.small_part, (int)grid, L.Gs, d.lds_off, d.G, qo);
        } else {
            const DimStruct& D = p->d[L.dims[o]];
            const bool timed_r = mode == kIter && kstats_enabled();
            if (timed_r) kstat_begin(kStatReduceRuns, s);
            hipLaunchKernelGGL((reduce_runs_kernel<CT>), dim3(stream_grid((size_t)d.G * CT, kBlock)), dim3(kBlock), 0,
                               s, (const double*)a.sd[o].runbuf, (const int*)D.run_order,
                               (const int64_t*)D.grp_run_off, d.G, qo);
            if (timed_r)
                kstat_end(kStatReduceRuns, s,
                          8.0 * CT * (double)(D.n_runs + d.G) + 4.0 * (double)D.n_runs + 8.0 * (double)(d.G + 1));
        }
        FEHIP_LAUNCH_CHECK("reduce kernel");
    }
    return 0;
}

struct Scratch {
    double* tab = nullptr;      // 7 table arrays
    double* runbuf[kMaxLarge] = {nullptr, nullptr, nullptr};
    double* small_part = nullptr;
    double* part = nullptr;     // reduction partials
    CgState* st = nullptr;
    CgState* st_host = nullptr;
    uint8_t* wpanel = nullptr;
    std::vector<void*> owned;
    ~Scratch() {
        for (void* v : owned) (void)hipFree(v);
        if (st_host) (void)hipHostFree(st_host);
    }
    template <typename T>
    int alloc(T** ptr, size_t count) {
        int e = dalloc(ptr, count);
        if (!e) owned.push_back((void*)*ptr);
        return e;
    }
};

constexpr int kCgBlocks = 512;

template <int CT>
int solve_tile(Panel* p, const Layout& L, Scratch& sc, const double* const* xcols, double* const* ocols, int ncols,
               const uint8_t* wpanel, const fehip_demean_opts& o, int to_work, TileStats& ts, hipStream_t s) {
    const int64_t S = L.S;
    const unsigned grid = apply_grid(p);
    ApplyArgs a{};
    a.seg_off = p->seg_off;
    a.GP = p->d[p->P].G;
    a.elim = L.elim;
    a.w = wpanel;
    a.nd = L.nd;
    for (int i = 0; i < L.nd; ++i) {
        a.sd[i] = L.sd[i];
        if (!L.sd[i].small) a.sd[i].runbuf = sc.runbuf[L.sd[i].large_slot];
    }
    a.Gs = L.Gs;
    for (int c = 0; c < 16; ++c) {
        a.xc[c] = c < ncols ? xcols[c] : nullptr;
        a.oc[c] = c < ncols ? ocols[c] : nullptr;
    }
    a.ncols = ncols;
    a.perm = p->perm;
    a.small_part = sc.small_part;
    a.tile_seg = p->tile_seg;
    a.n_tiles = p->n_tiles;
    a.long_segs = p->long_segs;
    a.n_long = p->n_long;
    a.to_work = to_work;
    double* xa = sc.tab;           // alpha
    double* r = xa + S * CT;
    double* z = r + S * CT;
    double* pp = z + S * CT;
    double* q = pp + S * CT;
    double* minv = q + S * CT;     // [S]
    double* wcnt = minv + S;       // [S]
    double* dg = wcnt + S;         // [S][2] diag + weight
    auto t0 = std::chrono::steady_clock::now();
    if (S > 0) {
        // Schur diagonal and weighted counts (CT = 2 pass per scattered dim)
        for (int od = 0; od < L.nd; ++od) {
            ApplyArgs ad = a;
            ad.diag_dim = od;
            // the diag pass writes 2 values per group: use table stride 2
            int e = apply_and_reduce<2>(p, L, ad, kDiag, dg, s);
            if (e) return e;
        }
        hipLaunchKernelGGL(make_minv_kernel, dim3(stream_grid(S, kBlock)), dim3(kBlock), 0, s, S, dg, minv, wcnt,
                           (const double*)dg);
        FEHIP_LAUNCH_CHECK("minv");
        // right-hand side b = D_o' W M_P x
        int e = apply_and_reduce<CT>(p, L, a, kData, r, s);
        if (e) return e;
        hipLaunchKernelGGL((cg_init_kernel<CT>), dim3(kCgBlocks), dim3(kBlock), 0, s, S, (const double*)r,
                           (const double*)minv, (const double*)wcnt, z, pp, xa, sc.part, sc.part + kCgBlocks * CT);
        hipLaunchKernelGGL((cg_start_scalars<CT>), dim3(1), dim3(64), 0, s, (const double*)sc.part,
                           (const double*)sc.part + kCgBlocks * CT, kCgBlocks, o.tol, ncols, sc.st);
        FEHIP_LAUNCH_CHECK("cg init");
        const int every = o.check_every > 0 ? o.check_every : 4;
        int it = 0;
        for (; it < o.max_iter; ++it) {
            if (it % every == 0) {
                FEHIP_CHECK(hipMemcpyAsync(sc.st_host, sc.st, sizeof(CgState), hipMemcpyDeviceToHost, s), "cg state");
                FEHIP_CHECK(hipStreamSynchronize(s), "cg sync");
                bool any = false;
                for (int c = 0; c < CT; ++c) any |= sc.st_host->active[c] != 0;
                if (!any) break;
            }
            a.p = pp;
            a.active = sc.st->active;
            e = apply_and_reduce<CT>(p, L, a, kIter, q, s);
            if (e) return e;
            hipLaunchKernelGGL((cg_pq_kernel<CT>), dim3(kCgBlocks), dim3(kBlock), 0, s, S, (const double*)pp,
                               (const double*)q, sc.part);
            hipLaunchKernelGGL((cg_alpha_scalars<CT>), dim3(1), dim3(64), 0, s, (const double*)sc.part, kCgBlocks,
                               sc.st);
            hipLaunchKernelGGL((cg_update_kernel<CT>), dim3(kCgBlocks), dim3(kBlock), 0, s, S, (const CgState*)sc.st,
                               (const double*)pp, (const double*)q, (const double*)minv, (const double*)wcnt, xa, r, z,
                               sc.part, sc.part + kCgBlocks * CT);
            hipLaunchKernelGGL((cg_beta_scalars<CT>), dim3(1), dim3(64), 0, s, (const double*)sc.part,
                               (const double*)sc.part + kCgBlocks * CT, kCgBlocks, o.tol, sc.st);
            hipLaunchKernelGGL((cg_pdir_kernel<CT>), dim3(stream_grid(S, kBlock)), dim3(kBlock), 0, s, S,
                               (const CgState*)sc.st, (const double*)z, pp);
            FEHIP_LAUNCH_CHECK("cg step");
        }
        FEHIP_CHECK(hipMemcpyAsync(sc.st_host, sc.st, sizeof(CgState), hipMemcpyDeviceToHost, s), "cg state");
        FEHIP_CHECK(hipStreamSynchronize(s), "cg sync");
        ts.iterations = 0;
        ts.converged = 1;
        ts.max_update = 0.0;
        for (int c = 0; c < ncols; ++c) {
            ts.iterations = std::max(ts.iterations, sc.st_host->iters[c]);
            if (sc.st_host->active[c]) ts.converged = 0;
            ts.max_update = std::max(ts.max_update, sc.st_host->maxn[c]);
        }
        ts.apply_launches = it;
    } else {
        ts.iterations = 0;
        ts.converged = 1;
        ts.max_update = 0.0;
        ts.apply_launches = 0;
    }
    // final pass over the columns
    {
        if (kstats_enabled()) kstat_begin(kStatFinalize, s);
        a.alpha = xa;
        hipLaunchKernelGGL((tile_fast_kernel<CT, kFinal>), dim3(tile_grid(p)), dim3(kBlock), tile_lds<CT>(L), s, a);
        FEHIP_LAUNCH_CHECK("finalize tile kernel");
        if (p->n_long > 0) {
            const size_t lds = sizeof(double) * (size_t)L.Gs * CT;
            hipLaunchKernelGGL((finalize_kernel<CT>), dim3(long_grid(p)), dim3(kBlock), lds, s, a, (const double*)xa,
                               to_work);
            FEHIP_LAUNCH_CHECK("finalize kernel");
        }
        if (kstats_enabled()) {
            double bytes = 16.0 * (double)p->n * ncols + 4.0 * (double)p->n * L.nd + 8.0 * (double)(a.GP + 1) +
                           (a.w ? (double)p->n : 0.0) + (p->perm ? 4.0 * (double)p->n : 0.0);
            for (int o = 0; o < L.nd; ++o) bytes += 8.0 * CT * (double)L.sd[o].G;
            kstat_end(kStatFinalize, s, bytes);
        }
    }
    FEHIP_CHECK(hipStreamSynchronize(s), "within sync");
    ts.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

}  // namespace

int within_transform(Panel* p, const double* const* xcols, int ncols, const uint8_t* w_caller,
                     const fehip_demean_opts& o, double* const* ocols, int to_work, WithinStats& ws, hipStream_t s) {
    unsigned mask = o.fe_mask ? o.fe_mask : ((1u << p->m) - 1u);
    mask &= (1u << p->m) - 1u;
    Layout L;
    int e = make_layout(p, mask, L);
    if (e) return e;
    ws = WithinStats();
    if (ncols <= 0) return 0;
    Scratch sc;
    const int64_t n = p->n;
    if (w_caller) {
        if ((e = sc.alloc(&sc.wpanel, (size_t)n))) return e;
        hipLaunchKernelGGL(weights_to_panel_kernel, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, s, w_caller,
                           p->perm, n, sc.wpanel);
        FEHIP_LAUNCH_CHECK("weights");
    }
    // column-tile width: widest of {16, 12, 9, 8, 6, 4, 3, 2, 1} whose small-dim LDS tables fit
    // next to the row tile (<= 28 KB of staged + accumulated tables)
    int ct_max = 16;
    {
        static const int widths[] = {16, 12, 9, 8, 6, 4, 3, 2, 1};
        ct_max = 0;
        for (int wdt : widths)
            if ((size_t)L.Gs * wdt * 16 <= 28 * 1024) {
                ct_max = wdt;
                break;
            }
        if (ct_max == 0) return set_error("small FE dims exceed the LDS budget (%d groups)", L.Gs);
    }
    const int64_t S = L.S;
    if ((e = sc.alloc(&sc.tab, (size_t)(S * 16 * 5 + S * 4 + 16)))) return e;
    for (int i = 0; i < L.nd; ++i)
        if (!L.sd[i].small) {
            const int64_t R = p->d[L.dims[i]].n_runs;
            if ((e = sc.alloc(&sc.runbuf[L.sd[i].large_slot], (size_t)(R * 16 + 1)))) return e;
        }
    const unsigned grid = apply_grid(p);
    if ((e = sc.alloc(&sc.small_part, (size_t)grid * (L.Gs > 0 ? L.Gs : 1) * 16))) return e;
    if ((e = sc.alloc(&sc.part, (size_t)kCgBlocks * 16 * 2))) return e;
    if ((e = sc.alloc(&sc.st, 1))) return e;
    FEHIP_CHECK(hipHostMalloc((void**)&sc.st_host, sizeof(CgState), hipHostMallocDefault), "pinned state");
    for (int c0 = 0; c0 < ncols; c0 += ct_max) {
        const int nc = std::min(ct_max, ncols - c0);
        static const int widths[] = {1, 2, 3, 4, 6, 8, 9, 12, 16};
        int ct = 16;
        for (int wdt : widths)
            if (wdt >= nc) {
                ct = wdt;
                break;
            }
        TileStats ts;
        const double* const* xt = xcols + c0;
        double* const* ot = ocols + c0;
        switch (ct) {
            case 1: e = solve_tile<1>(p, L, sc, xt, ot, nc, sc.wpanel, o, to_work, ts, s); break;
            case 2: e = solve_tile<2>(p, L, sc, xt, ot, nc, sc.wpanel, o, to_work, ts, s); break;
            case 3: e = solve_tile<3>(p, L, sc, xt, ot, nc, sc.wpanel, o, to_work, ts, s); break;
            case 4: e = solve_tile<4>(p, L, sc, xt, ot, nc, sc.wpanel, o, to_work, ts, s); break;
            case 6: e = solve_tile<6>(p, L, sc, xt, ot, nc, sc.wpanel, o, to_work, ts, s); break;
            case 8: e = solve_tile<8>(p, L, sc, xt, ot, nc, sc.wpanel, o, to_work, ts, s); break;
            case 9: e = solve_tile<9>(p, L, sc, xt, ot, nc, sc.wpanel, o, to_work, ts, s); break;
            case 12: e = solve_tile<12>(p, L, sc, xt, ot, nc, sc.wpanel, o, to_work, ts, s); break;
            default: e = solve_tile<16>(p, L, sc, xt, ot, nc, sc.wpanel, o, to_work, ts, s); break;
        }
        if (e) return e;
        ws.iterations = std::max(ws.iterations, ts.iterations);
        ws.converged = ws.converged && ts.converged;
        ws.max_update = std::max(ws.max_update, ts.max_update);
        ws.seconds += ts.seconds;
        ws.apply_launches += ts.apply_launches;
        // algorithmic bytes: ids (+weights) per apply, columns read twice and written once
        const int64_t id_bytes = n * 4 * (L.nd + (L.elim ? 0 : 0)) + (int64_t)(p->d[p->P].G + 1) * 8 +
                                 (sc.wpanel ? n : 0);
        ws.bytes += ts.apply_launches * id_bytes + (int64_t)3 * n * 8 * nc;
    }
    return 0;
}

}  // namespace fehip

using namespace fehip;

extern "C" int fehip_demean(fehip_panel* panel, double* d_cols, int n_cols, int64_t ld, const uint8_t* d_weights,
                            const fehip_demean_opts* opts, fehip_demean_stats* stats) {
    (void)hipGetLastError();  // start from a clean error state
    if (!panel || !opts) return set_error("fehip_demean: null argument");
    if (n_cols < 0 || (n_cols > 0 && (!d_cols || ld < ((Panel*)panel)->n))) return set_error("fehip_demean: bad columns");
    Panel* p = reinterpret_cast<Panel*>(panel);
    std::vector<double*> cols(n_cols);
    for (int c = 0; c < n_cols; ++c) cols[c] = d_cols + (int64_t)c * ld;
    WithinStats ws;
    int e = within_transform(p, cols.data(), n_cols, d_weights, *opts, cols.data(), 0, ws, 0);
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
