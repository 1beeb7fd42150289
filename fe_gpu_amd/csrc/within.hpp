// Within-transform in FE-table space (see within.hip).
#pragma once

#include <vector>

#include "panel.hpp"

namespace fehip {

struct TileStats {
    int iterations = 0;
    int converged = 1;
    double max_update = 0.0;
    double seconds = 0.0;
    int apply_launches = 0;
    int col_iters[16] = {0};
    int col_converged[16] = {0};
    double col_max_update[16] = {0};
};

struct WithinStats {
    explicit WithinStats(int ncols = 0) : col_iters(ncols, 0), col_converged(ncols, 1), col_max_update(ncols, 0.0) {}
    int iterations = 0;
    int converged = 1;
    double max_update = 0.0;
    double seconds = 0.0;
    int64_t apply_launches = 0;
    int64_t bytes = 0;
    std::vector<int> col_iters;      // per column: CG iterations
    std::vector<int> col_converged;  // per column: criterion met
    std::vector<double> col_max_update;  // per column: last max |r_g / n_g|
};

// Columns of one solve: input and output device pointers and the request slot of each column.
struct WithinCols {
    explicit WithinCols(int n) : x(n), out(n), slot(n, 0) {}
    std::vector<const double*> x;
    std::vector<double*> out;
    std::vector<int> slot;
};

// Panel-order row mask words from up to 32 caller-order 0/1 weight arrays (nullptr = every row).
int build_mask_words(Panel* p, const uint8_t* const* w_caller, int nslots, uint32_t* wm, hipStream_t s);

// M_F on ncols columns (caller row order).  wm: panel-order mask words (nullptr: unweighted, one
// slot).  to_work = 1 writes the result in panel order (zeros on rows outside the column's slot);
// to_work = 0 writes back in caller order (rows outside the slot untouched).
int within_transform(Panel* p, const WithinCols& wc, int ncols, const uint32_t* wm, int nslots,
                     const fehip_demean_opts& o, int to_work, WithinStats& ws, hipStream_t s);

}  // namespace fehip
