// tgtp_fast.hip -- symmetrization of the FAST modes' FP64 row sums.
//
// preds holds, per CSR position, the FP64 row sum of its source's
// diffusion and walk terms; the score of an edge is the sum of both
// directions' row sums (aesc.h:532-550: preds[pos] += preds_old[rev(pos)]),
// rounded to float once.  src_of[t] is the CSR position whose row sum is
// added to t's (the first occurrence of the reverse edge, as the
// reference's lower_bound finds it), or -1.
#include <algorithm>

#include "common.hpp"
#include "tgtp_fast.hpp"

namespace {

__global__ void __launch_bounds__(256)
symmetrize_d_kernel(const double* __restrict__ preds, const int32_t* __restrict__ src_of,
                    uint64_t m, float* __restrict__ out) {
  for (uint64_t t = blockIdx.x * 256ull + threadIdx.x; t < m; t += gridDim.x * 256ull) {
    const int32_t s = src_of[t];
    out[t] = static_cast<float>(preds[t] + (s >= 0 ? preds[s] : 0.0));
  }
}

}  // namespace

int launch_symmetrize_d(tgtp_ctx* ctx, const double* preds, const int32_t* src_of, uint64_t m,
                        float* out) {
  const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>(4096, (m + 255) / 256));
  symmetrize_d_kernel<<<static_cast<uint32_t>(blocks), 256, 0, ctx->stream>>>(preds, src_of, m, out);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}
