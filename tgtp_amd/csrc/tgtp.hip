// tgtp.hip -- TGT+ entry points (tau bounds, per-source diffusion, walks,
// symmetrization).
#include "common.hpp"

extern "C" int tgtp_taus(tgtp_ctx* ctx, const tgtp_graph*, const tgtp_eigens*, float, int32_t,
                         int32_t*, tgtp_stats*) {
  return set_error(ctx, TGTP_E_UNSUPPORTED, "tgtp_taus: not built yet");
}

extern "C" int tgtp_score(tgtp_ctx* ctx, const tgtp_graph*, const tgtp_eigens*,
                          const tgtp_params*, float*, tgtp_stats*) {
  return set_error(ctx, TGTP_E_UNSUPPORTED, "tgtp_score: not built yet");
}

extern "C" int tgtp_score_dev(tgtp_ctx* ctx, const tgtp_dgraph*, const tgtp_eigens*,
                              const tgtp_params*, float*, tgtp_stats*) {
  return set_error(ctx, TGTP_E_UNSUPPORTED, "tgtp_score_dev: not built yet");
}
