// Within-transform in FE-table space (see within.hip).
#pragma once

#include "panel.hpp"

namespace fehip {

struct TileStats {
    int iterations = 0;
    int converged = 1;
    double max_update = 0.0;
    double seconds = 0.0;
    int apply_launches = 0;
};

struct WithinStats {
    int iterations = 0;
    int converged = 1;
    double max_update = 0.0;
    double seconds = 0.0;
    int64_t apply_launches = 0;
    int64_t bytes = 0;
};

// M_F on ncols columns given as device pointers (caller row order).  to_work = 1 writes the
// result in panel order into ocols (zeros on rows with weight 0); to_work = 0 writes back in
// caller order (rows with weight 0 untouched).
int within_transform(Panel* p, const double* const* xcols, int ncols, const uint8_t* w_caller,
                     const fehip_demean_opts& o, double* const* ocols, int to_work, WithinStats& ws, hipStream_t s);

}  // namespace fehip
