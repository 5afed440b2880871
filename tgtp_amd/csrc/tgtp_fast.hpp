// tgtp_fast.hpp -- launch interface of the FAST_FP32 / FAST_BF16 per-source
// pipeline (tgtp_fast.hip) and of the tau projection (taus_fast.hip).
#pragma once

#include <cstddef>
#include <cstdint>

#include "common.hpp"

struct FastArgs {
  const uint2* rowd;
  const uint32_t* adj;
  const float* degf;
  const int32_t* taus;
  double* preds_d;    // [m] FP64 row sums
  uint64_t* nr_pos;   // [m]
  uint32_t* qbuf;     // [slots][2n]
  float* pv_g;        // [slots][2n + n/32] when the state does not fit LDS
  uint32_t* next_src;
  uint32_t* status;   // bit 0: walk count beyond 2^36
  uint32_t n;
  int32_t gamma;
  double eps;
  double logmd;
  int32_t* tr_ell;
  float* tr_emax;
  float* tr_gmax;
  float* tr_gmin;
  uint32_t* tr_qsize;
  float* tr_x;
};

size_t fast_lds_bytes(uint64_t n);
bool fast_fits_lds(uint64_t n);
int launch_fast_sources(tgtp_ctx* ctx, const FastArgs& A, uint32_t slots, bool use_lds);
int launch_symmetrize_d(tgtp_ctx* ctx, const double* preds, const int32_t* src_of, uint64_t m,
                        float* out);
int fast_slots_per_cu(bool use_lds, uint64_t n);
