// tgtp_internal.hpp -- pieces of the tgtp pipeline shared between the
// faithful path (tgtp.hip) and the fast / batched path (tgtp_batch.hip).
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "common.hpp"

// Host-side state kept with a device graph once tgtp has run on it.
struct TgtpPrep {
  std::vector<uint32_t> h_idx, h_adj;
  int32_t* d_src_of = nullptr;  // symmetrization source per CSR position
  void* fast_plan = nullptr;    // cached FAST plan (tgtp_batch.hip)
  void (*fast_plan_free)(void*) = nullptr;
};

// Per-stage outputs of tgtp_score_trace (any pointer may be null).
struct TraceOut {
  int32_t* taus;
  uint64_t* nr;
  float* walk_sum;
  float* preds_row;
  int32_t* ell;
  float* edge_max;
  float* gmax;
  float* gmin;
  uint32_t* q_size;
};

// tgtp_params.mode = TGTP_MODE_* | optional TGTP_TAU_REFERENCE (FAST) or
// TGTP_WALK2_MT19937 (FAITHFUL)
inline int32_t tgtp_mode_base(int32_t mode) { return mode & 0xff; }
inline bool tgtp_mode_valid(int32_t mode) {
  const int32_t b = tgtp_mode_base(mode);
  if ((mode & TGTP_WALK2_MT19937) && b != TGTP_MODE_FAITHFUL) return false;
  return (mode & ~(0xff | TGTP_TAU_REFERENCE | TGTP_WALK2_MT19937)) == 0 &&
         (b == TGTP_MODE_FAITHFUL || b == TGTP_MODE_FAST_FP32 || b == TGTP_MODE_FAST_BF16);
}

int tgtp_ensure_prep(tgtp_ctx* ctx, tgtp_dgraph* dg);
int tgtp_ensure_eigens(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e);
int tgtp_check_degrees(tgtp_ctx* ctx, tgtp_dgraph* dg);
int tgtp_run_taus_faithful(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e, float eps_half,
                           int32_t* d_taus, uint64_t* n_fixups);
int launch_fill_i32(tgtp_ctx* ctx, int32_t* p, uint64_t n, int32_t v);
int launch_fill_words(tgtp_ctx* ctx, uint32_t* p, uint32_t n);
// per-position tau0 (initial bound) maximum of each graph of a union CSR
int launch_tau0_per_graph(tgtp_ctx* ctx, const uint32_t* idx, const uint32_t* adj, uint64_t m,
                          uint32_t n, const uint32_t* voff, uint32_t n_graphs,
                          const double* g_params /* [n_graphs][3]: eps, sq1m_1, b_1 */,
                          uint32_t* g_tau0max);
double tgtp_now_ms();

// FAST modes (tgtp_batch.hip)
int tgtp_run_score_fast(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e,
                        const tgtp_params* p, float* scores_out, tgtp_stats* stats,
                        const TraceOut* tr);
int tgtp_run_score_fast_sources(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e,
                                const tgtp_params* p,
                                const std::vector<std::pair<uint32_t, uint32_t>>& ranges,
                                float* scores_all, float* rows_all, tgtp_stats* stats);
int tgtp_run_taus_fast(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e, float eps_half,
                       int32_t mode, int32_t* d_taus);
