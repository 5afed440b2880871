// tgtp_fast.hpp -- symmetrization of the FAST modes' FP64 row sums
// (tgtp_fast.hip).
#pragma once

#include <cstdint>

#include "common.hpp"

// out[t] = float(preds[t] + preds[src_of[t]]) (src_of[t] < 0: no reverse
// entry), the FP64 form of aesc.h:532-550.
int launch_symmetrize_d(tgtp_ctx* ctx, const double* preds, const int32_t* src_of, uint64_t m,
                        float* out);
