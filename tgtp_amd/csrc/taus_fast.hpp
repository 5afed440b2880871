// taus_fast.hpp -- launch interface of the MFMA tau projection (taus_fast.hip).
// Works on a single graph or on the disjoint union of a batch: 16-edge tiles
// never span two graphs, and each tile reads its graph's column table and
// scalars.
#pragma once

#include <cstdint>

#include "common.hpp"

struct TauGraph {          // per graph of the batch
  double sq1m_1, b_1;      // 1 - lambda_1^2, log(1/|lambda_1|)
  double sq1m_w, b_w;      // same for lambda_{omega-1}
  float m;                 // CSR entries as float (the reference divides by float(m))
  float eps;               // float(epsilon / 2)
  uint32_t wt_row0;        // first row of this graph's column table in Wt
  uint32_t n_col_tiles;
};

struct TauFastArgs {
  const uint32_t* idx;
  const uint32_t* adj;
  const uint32_t* und;        // [n_tiles*16] CSR positions (0xFFFFFFFF = padding)
  const uint32_t* tile_graph; // [n_tiles] graph of each tile
  const TauGraph* graphs;     // [n_graphs]
  uint64_t n_tiles;
  const float* X;             // [N*omega] FP32 eigenvectors (FAST_FP32)
  const __bf16* Xb;           // [N*omega] BF16 eigenvectors (FAST_BF16)
  const float* Wt;            // [rows][omega_pad] column tables, FP32
  const __bf16* Wtb;          // same, BF16
  uint32_t n;                 // total vertices
  int32_t omega;
  uint32_t omega_pad;         // multiple of 32, <= 128
  uint32_t bounded;           // 1: skip refinement steps whose top-omega tail reaches eps/2
  int32_t* taus;
};

int launch_taus_fast(tgtp_ctx* ctx, const TauFastArgs& A, bool bf16);
int launch_to_bf16(tgtp_ctx* ctx, const float* x, uint64_t count, void* y);
