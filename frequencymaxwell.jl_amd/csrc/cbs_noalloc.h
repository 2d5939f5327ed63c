// The allocation-free operator apply (cbs_operator_apply_noalloc): declarations
// shared by cbs_api.hip (C ABI, context set-up) and cbs_noalloc.hip.
#pragma once

#include "cbs_engine.h"

namespace cbs {

// Resolve the address of this device's copy of the scratch array (once per
// device, with that device current).  Called by cbs_ctx_create.
void noalloc_init(int device);

// y <- alpha (x - G_avg(V x)) + beta y on device arrays x, y (3 nvox complex in
// the problem's working precision: FP64 for CBS_FP64 problems, else FP32),
// with the spectrum workspace and the launch tables in the scratch array.
// Allocates nothing; synchronises the context's stream before returning.
void operator_apply_noalloc(Problem& p, void* y, const void* x, double ar, double ai,
                            double br, double bi);

}  // namespace cbs
