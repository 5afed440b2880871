// taus_fast.hpp -- launch interface of the MFMA tau projection (taus_fast.hip).
#pragma once

#include <cstdint>

#include "common.hpp"

struct TauFastArgs {
  const uint32_t* idx;
  const uint32_t* adj;
  const uint32_t* und;   // [n_und] CSR positions p with adj[p] >= row(p)
  uint64_t n_und;
  const float* X;        // [n*omega] FP32 eigenvectors (FAST_FP32)
  const __bf16* Xb;      // [n*omega] BF16 eigenvectors (FAST_BF16)
  const float* Wt;       // [n_col_tiles*16][omega_pad] column table, FP32
  const __bf16* Wtb;     // same, BF16
  uint64_t m;
  uint32_t n;
  int32_t omega;
  uint32_t omega_pad;    // multiple of 32, <= 128
  uint32_t n_col_tiles;
  float eps;             // float(epsilon / 2)
  double sq1m_1, b_1;    // 1 - lambda_1^2, log(1/|lambda_1|)
  double sq1m_w, b_w;    // same for lambda_{omega-1}
  int32_t* taus;
};

int launch_taus_fast(tgtp_ctx* ctx, const TauFastArgs& A, bool bf16);
int launch_to_bf16(tgtp_ctx* ctx, const float* x, uint64_t count, void* y);
