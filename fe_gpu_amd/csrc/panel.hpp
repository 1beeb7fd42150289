// Device-resident panel: FE ids re-ordered by the primary (largest) FE dim, with the segment,
// run and cluster-partition structures every coarse kernel reads.
#pragma once

#include <map>
#include <mutex>
#include <vector>

#include "../../include/fehip.h"
#include "common.hpp"

namespace fehip {

// Small dims keep their whole group table in LDS; larger dims scatter through runs.
constexpr int kSmallMaxGroups = 1024;
// Rows per tile: a tile is a run of whole primary segments holding at most kTileRows rows,
// processed by one 256-thread workgroup with one row per thread.
constexpr int kTileRows = 256;
// Rows per wave tile: one row per lane of a 64-wide wavefront.
constexpr int kWaveRows = 64;

struct DimStruct {
    int G = 0;                 // groups (max id)
    int small = 0;             // table fits LDS
    int* ids = nullptr;        // panel order, 0-based
    int* counts = nullptr;     // rows per group (unweighted)
    // runs: maximal stretches of equal id inside one primary segment (non-primary dims only)
    int64_t n_runs = 0;
    int64_t* seg_run_off = nullptr;  // [G_P + 1] first run of each primary segment
    int* run_order = nullptr;        // run indices sorted by group
    int64_t* grp_run_off = nullptr;  // [G + 1] offsets into run_order
    int* run_dest = nullptr;         // position of each run in group order (inverse of run_order)
    int* row_rdest = nullptr;        // [n] per row: run_dest of the run it ends, -1 inside a run
    int* run_gid = nullptr;          // [n_runs] group of each run (runs in panel order)
    int* run_len = nullptr;          // [n_runs] rows of each run; bit 31 set on a segment's first run
    // run tiles (built on first use): whole primary segments packed into at most 64 runs each
    // (first run | runs << 32), and the segments of more runs, listed apart
    int64_t n_rtiles = -1;
    int64_t* rtile_rec = nullptr;
    int64_t* rtile_seg = nullptr;    // first primary segment of each run tile
    int64_t n_rlong = 0;
    int64_t* rlong_segs = nullptr;
    // the runs in group order (run_order): each one's primary segment and length (built with the run tiles)
    int* grp_seg = nullptr;
    int* grp_len = nullptr;
};

// The couplings of one large non-primary dim D with the small absorbed dims (all of them, Gs groups)
// in the FE-space operator S = D_o' M_P D_o of an unweighted solve: B = D_D' M_P D_s ([G_D][Gs]) and
// C = D_s' M_P D_s ([Gs][Gs]), so that S p = (D_D' M_P D_D p_D + B p_s, B' p_D + C p_s) with the first
// term formed from D's runs alone.  Built once per absorbed-dims mask.
struct RunCoupling {
    // 0: S p = (A p_D + B p_s, B' p_D + C p_s) with A from D's runs in the primary segments (primary absorbed);
    // 1: the primary not absorbed, so A = D_D' D_D is the diagonal of D's group counts;
    // 2: no large dim, S = C
    int kind = 0;
    int o_large = -1;               // layout index of D (-1: none)
    int Gs = 0;
    int* small_rows = nullptr;      // [Gs] table rows of the small groups
    double* B = nullptr;            // [G_D][Gs]
    float* B32 = nullptr;           // B in FP32, for the FP32 operator's products
    double* C = nullptr;            // [Gs][Gs]
    // a solve with a row subsample: its runs' lengths counted over the subsample (run order, with the
    // segment-head bit, and group order); nullptr: the panel's run lengths
    const int* run_len = nullptr;
    const int* grp_len = nullptr;
};

// Row partition of a cluster subset: rows of cluster c are order[off[c] .. off[c+1]).
struct Partition {
    int64_t n_clusters = 0;
    int* order = nullptr;      // nullptr => identity (panel order)
    int64_t* off = nullptr;    // [n_clusters + 1]
    int* keys = nullptr;       // cluster of each sorted position (grouped, nondecreasing)
    int* pos = nullptr;        // sorted position of each row (inverse of order), built on first use
};

struct Panel {
    std::recursive_mutex mu;   // guards the structures built lazily on first use (partitions, caches)
    int device = 0;            // HIP device holding the panel
    int64_t n = 0;
    int m = 0;
    int P = 0;                 // primary dim
    int sorted_input = 1;
    int max_seg = 0;
    int* perm = nullptr;       // panel row -> caller row (nullptr if identity)
    int64_t* seg_off = nullptr;  // [G_P + 1]
    int64_t n_tiles = 0;
    int64_t* tile_seg = nullptr;   // [2 * n_tiles]: first and one-past-last segment of each tile
    int64_t n_long = 0;
    int64_t* long_segs = nullptr;  // segments longer than kTileRows (handled one thread per segment)
    // the operator kernels' partition of the non-long segments: wave tiles (whole segments of at most
    // kWaveRows rows each, packed into at most kWaveRows rows, one wavefront per tile) and block tiles
    // (whole segments longer than kWaveRows, at most kTileRows rows per tile)
    int64_t n_wtiles = 0;
    int64_t* wtile_seg = nullptr;  // [2 * n_wtiles]
    int64_t* wtile_rec = nullptr;  // [n_wtiles] first row | rows << 32
    int64_t n_mtiles = 0;
    int64_t* mtile_seg = nullptr;  // [2 * n_mtiles]
    DimStruct d[FEHIP_MAX_FE];
    std::map<unsigned, Partition> partitions;  // key: bit mask of 0-based dims
    // subsets containing the primary: 2 = every row its own cluster, 1 = each cluster is one run inside
    // its segment, 0 = runs of one segment may share a cluster
    std::map<unsigned, int> nest_modes;
    // Schur diagonal and group counts of unweighted solves, per absorbed-dims mask ([S][2])
    std::map<unsigned, double*> diag_cache;
    // run-level operator couplings of unweighted solves, per absorbed-dims mask
    std::map<unsigned, RunCoupling> run_couplings;
    // measured time of one operator product per column-tile width (index into the width list), per
    // absorbed-dims mask; filled by the within transform's tuner on first use
    std::map<unsigned, std::vector<float>> width_ms;
    // FP32 copies of registered FP64 columns (caller row order), used by FP32-storage requests
    struct RegCols {
        const double* base;
        int64_t ld;
        int ncols;
        float* copy;  // [ncols][n]
    };
    std::vector<RegCols> reg_cols;
    int64_t bytes = 0;
};

// The FP32 copy of FP64 column x if x lies in a registered block, else nullptr.
const float* panel_col32(Panel* p, const double* x);

int panel_partition(Panel* p, unsigned dims_mask, const Partition** out, hipStream_t s);
// The partition's sorted position of every row (nullptr for a partition in panel order), cached.
int partition_positions(Panel* p, unsigned dims_mask, const int** pos, hipStream_t s);

// gemm on the default stream (linalg.hip); C = alpha op(A) op(B) + beta C
int gemm_impl(bool tA, bool tB, int m, int n, long long k, double alpha, const double* A, long long ldA,
              const double* B, long long ldB, double beta, double* C, long long ldC, bool upper_only);

int build_intersection_ids(const int* const* dims, const unsigned long long* strides, int nd, long long n, int* out,
                           int* n_clusters, hipStream_t stream);

// device allocation helper that tracks panel bytes
template <typename T>
int dalloc(T** p, size_t count, int64_t* tally = nullptr) {
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = sizeof(T);
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_hip_error(e, "hipMalloc");
    }
    if (tally) *tally += (int64_t)bytes;
    return 0;
}

// per-call temporaries from the caching allocator
template <typename T>
int talloc(T** p, size_t count) {
    *p = (T*)cache_alloc(count * sizeof(T));
    return *p ? 0 : set_error("device allocation of %zu bytes failed", count * sizeof(T));
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)locked_free((void*)p);
    p = nullptr;
}

}  // namespace fehip
