// Per-kernel timing with HIP events on the launch stream (opt-in; used by the benchmark to
// report the dominant kernel's average duration and algorithmic bytes per launch).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fehip {

// apply_iter32: the FP32 operator product of the mixed-precision solve; apply_resid: its FP64 products
// for the true residual b - S x; couplings: the one-time operator products that build the run-level
// operator's couplings with the small dims
enum KStatId {
    kStatApplyIter = 0, kStatApplyData, kStatFinalize, kStatReduceRuns, kStatGram, kStatSegsum, kStatApplyIter32,
    kStatApplyResid, kStatCouplings, kStatCount
};

bool kstats_enabled();
// Record a start event before a launch and the stop event after it; bytes/flops are the
// algorithmic work of that launch.  No-ops when disabled.
void kstat_begin(int id, hipStream_t s);
void kstat_end(int id, hipStream_t s, double bytes, double flops = 0.0);

}  // namespace fehip
