// Whole-regression entry used by the batch API and the request server (see estimate.hip).
#pragma once

#include "panel.hpp"

namespace fehip {

// Estimate n_req requests on one panel: requests with the same FE set share the within-transform
// sweeps (up to 32 per sweep, chunked by free device memory), then each takes its own normal-equation
// and sandwich step.  Runs on the calling thread's current device and the null stream.
int estimate_requests(Panel* p, const fehip_request* reqs, int n_req);

}  // namespace fehip
