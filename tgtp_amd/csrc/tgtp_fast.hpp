// tgtp_fast.hpp -- launch interface of the FAST_FP32 / FAST_BF16 per-source
// pipeline (tgtp_fast.hip).
//
// The kernels run on a CSR that may be the disjoint union of several graphs
// (the batched form): graph k owns vertices [voff[k], voff[k+1]) and every
// per-graph quantity of AESC::tgtp (n, m/delta, epsilon, gamma) is looked up
// from the source's graph.  A single graph is the union of one.
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
  uint32_t* qbuf;     // [slots][2 * qcap]
  float* pv_g;        // [slots][2 * slot_n + slot_n/32] (graphs above lds_n)
  uint32_t* next_src;
  uint32_t* status;   // bit 0: walk count beyond 2^36
  const uint32_t* voff;    // [n_graphs + 1]
  uint32_t n_graphs;
  const double* g_eps;     // [n_graphs]
  const double* g_logmd;   // [n_graphs] log(m_k / delta_k)
  const int32_t* g_gamma;  // [n_graphs]
  const uint32_t* work;    // processing order of sources (null: 0 .. n_work-1)
  uint32_t n_work;
  uint32_t lds_n;    // graphs with at most lds_n vertices keep pvec/bitmap in LDS
  uint32_t slot_n;   // HBM state capacity per slot (vertices)
  uint32_t qcap;     // frontier queue capacity per slot (vertices)
  uint32_t light;    // degree up to which one thread handles a vertex
  uint32_t small_bq; // walk positions with at most this many bouquets: one 8-lane group
  int32_t* tr_ell;
  float* tr_emax;
  float* tr_gmax;
  float* tr_gmin;
  uint32_t* tr_qsize;
  float* tr_x;
};

size_t fast_lds_bytes(uint64_t n);
bool fast_fits_lds(uint64_t n);
// floats of per-source state for a graph of n vertices: pvec, next pvec, bitmap
__host__ __device__ inline uint64_t fast_slot_floats(uint64_t n) { return 2 * n + (n + 31) / 32; }
int launch_fast_sources(tgtp_ctx* ctx, const FastArgs& A, uint32_t slots);
int fast_slots_per_cu(uint32_t lds_n);
int launch_symmetrize_d(tgtp_ctx* ctx, const double* preds, const int32_t* src_of, uint64_t m,
                        float* out);
