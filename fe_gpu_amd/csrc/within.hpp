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
    int refine_launches = 0;  // FP64 true-residual products of the mixed-precision solve
    int col_iters[16] = {0};
    int col_converged[16] = {0};
    double col_max_update[16] = {0};
};

struct WithinStats {
    explicit WithinStats(int ncols = 0)
        : col_iters(ncols, 0), col_converged(ncols, 1), col_max_update(ncols, 0.0), col_tile(ncols, 0),
          col_refines(ncols, 0) {}
    int iterations = 0;
    int converged = 1;
    double max_update = 0.0;
    double seconds = 0.0;
    int64_t apply_launches = 0;
    int64_t bytes = 0;
    std::vector<int> col_iters;      // per column: CG iterations
    std::vector<int> col_converged;  // per column: criterion met
    std::vector<double> col_max_update;  // per column: last max |r_g / n_g|
    std::vector<int> col_tile;       // per column: column tile it was solved in
    std::vector<int> col_refines;    // per column: FP64 true-residual products (FP32-operator solve)
};

// Right-hand sides formed once for the solves of several unweighted FE sets: b ([S_u][ld]) holds
// D_o' M_P x (primary absorbed) or D_o' x over the union of those sets' dims (within_rhs); a solve takes
// the blocks of its own dims.  dim_off: table offset of each panel dim in b's layout (-1: not in it).
struct SharedRhs {
    const double* b = nullptr;
    int ld = 0;
    int64_t dim_off[FEHIP_MAX_FE];
};

// Columns of one solve: input and output device pointers and the request slot of each column.
// out32 (optional, per column) takes the result in FP32 instead of out; alpha (optional, per
// column) receives the solved FE tables ([S] doubles, S = within_table_size).
struct WithinCols {
    explicit WithinCols(int n) : x(n), out(n), out32(n, nullptr), alpha(n, nullptr), slot(n, 0) {}
    std::vector<const double*> x;
    std::vector<const float*> x32;  // optional (empty, or one per column): FP32 copies of x, or nullptr
    std::vector<double*> out;
    std::vector<float*> out32;
    std::vector<double*> alpha;
    std::vector<int> slot;
    // optional device matrix (ncols x ncols, column-major, leading dim gram_ld) that receives the FP64
    // cross products of the transformed columns (rows outside a column's slot count as 0), formed in
    // the final pass before any rounding to FP32 storage.  Pairs in different column tiles are left
    // untouched (WithinStats::col_tile); with at most 16 columns the solve uses one tile.
    double* gram = nullptr;
    int64_t gram_ld = 0;
    // optional: the right-hand sides come from a shared computation (column rhs_col[c] of rhs->b)
    const SharedRhs* rhs = nullptr;
    std::vector<int> rhs_col;
};

// Table offset of each panel dim in the layout of fe_mask (-1: not scattered there).
int within_dim_offsets(Panel* p, unsigned fe_mask, int64_t* off);

// The right-hand sides of nx <= 16 caller-order columns over fe_mask (unweighted), into b ([S][*ld],
// *ld the column-tile width used, at most 16; b holds S * 16 doubles).  x32 (optional): FP32 copies of the
// columns, read instead when every column has one.
int within_rhs(Panel* p, unsigned fe_mask, const double* const* x, const float* const* x32, int nx, double* b, int* ld,
               hipStream_t s);


// Widest column tile of a solve over fe_mask (columns that one tile, hence one set of cross products,
// can hold); 0 if the small dims do not fit the LDS budget.
int within_tile_cap(Panel* p, unsigned fe_mask);

// Size of the FE-table space of a solve over fe_mask (groups of the absorbed non-primary dims).
int64_t within_table_size(Panel* p, unsigned fe_mask);

// One FP64 sweep M_P(x - D_o alpha) of a single column x (caller order) with given tables alpha,
// into out (panel order, zeros on rows outside the slot).  The tables of a linear combination of
// solved columns are the same combination of their tables, so this forms M_F of a combination
// (the regression residual) exactly without another solve.
int within_sweep_with_tables(Panel* p, unsigned fe_mask, const double* x, const double* alpha, const uint32_t* wm,
                             int slot, double* out, hipStream_t s);

// Panel-order row mask words from up to 32 caller-order 0/1 weight arrays (nullptr = every row).
int build_mask_words(Panel* p, const uint8_t* const* w_caller, int nslots, uint32_t* wm, hipStream_t s);

// M_F on ncols columns (caller row order).  wm: panel-order mask words (nullptr: unweighted, one
// slot).  to_work = 1 writes the result in panel order (zeros on rows outside the column's slot);
// to_work = 0 writes back in caller order (rows outside the slot untouched).
int within_transform(Panel* p, const WithinCols& wc, int ncols, const uint32_t* wm, int nslots,
                     const fehip_demean_opts& o, int to_work, WithinStats& ws, hipStream_t s);

}  // namespace fehip
