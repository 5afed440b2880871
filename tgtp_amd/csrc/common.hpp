// common.hpp -- shared pieces of the HIP library: the context, device CSR,
// status plumbing and the integer mixers used by every walk kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "tgtp_hip.h"

// ---- integer mixers --------------------------------------------------------
// MurmurHash3 fmix32 (aesc.h:31-41).  Applied to the current vertex id; the
// result is xored into the walk's grid secret to pick the next neighbour.
__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// SplitMix32 16/15/15 (aesc.h:23-30): the fallback generator, one state per
// walk, seeded with the walk's initial grid secret.
__host__ __device__ __forceinline__ uint32_t splitmix32(uint32_t& state) {
  state += 0x9e3779b9u;
  uint32_t z = state ^ (state >> 16);
  z *= 0x21f0aaadu;
  z ^= z >> 15;
  z *= 0x735a2d97u;
  return z ^ (z >> 15);
}

// ---- device-resident CSR ----------------------------------------------------
// rowd[v] = {idx[v], deg(v)}: one aligned 8-byte load per hop instead of two
// dependent 4-byte loads of idx[v] and idx[v+1].
struct tgtp_dgraph {
  int device = -1;
  uint64_t n = 0, m = 0;
  uint32_t* idx = nullptr;  // [n+1]
  uint32_t* adj = nullptr;  // [m]
  uint2* rowd = nullptr;    // [n]
  uint32_t* rev = nullptr;  // [m], built on first tgtp use
  float* degf = nullptr;    // [n], float(deg)
  uint64_t max_deg = 0;
  // packed walk layout {id | row start | degree} per CSR entry, built by the
  // first walk call when the three fields fit in 64 bits
  uint64_t* walk_adj = nullptr;  // [m]
  uint32_t id_bits = 0, start_bits = 0;
  // degree share of each cold vertex range of the bucketed walk counting,
  // computed once per (hot, bshift) from the row degrees
  std::vector<double> walk_mass;
  uint32_t walk_mass_hot = 0, walk_mass_bshift = 0;
  // vertex partitions of the partitioned walk engine (saba_part.hip): bounds
  // on the host, bounds and the coarse lookup table on the device
  std::vector<uint32_t> part_plo;
  uint32_t part_vmax = 0, part_want = 0, part_tshift = 0, part_balance = 0;
  void* part_dev = nullptr;
  // adjacency with each neighbour's partition in bits 24-31 (n <= 2^24),
  // valid for the current partitions when part_tagged
  uint32_t* part_adj = nullptr;
  bool part_tagged = false;
  // tgtp host-side state (host CSR copy, symmetrization map), built lazily
  void* prep = nullptr;
  void (*prep_free)(void*) = nullptr;
  // cached eigenpairs: the caller's vectors / values pointers and omega of
  // the last upload, and a count of uploads (plans built from them match it)
  const void* eig_key = nullptr;
  const void* eig_vals = nullptr;
  int32_t eig_omega = 0;
  uint64_t eig_gen = 0;
  double* lam = nullptr;    // [omega]
  float* X = nullptr;       // [n*omega]
};

// ---- context -----------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct tgtp_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};
  std::string err;
  DevBuf bufs[24];
  void* scratch = nullptr;  // this device's copy of capi.hip's static scratch (no-allocation entry)
};

// Scratch slots in tgtp_ctx::bufs.
enum BufSlot {
  kBufCounters = 0,
  kBufSums = 1,
  kBufTaus = 2,
  kBufPreds = 3,
  kBufPvec = 4,
  kBufQueue = 5,
  kBufSrcInfo = 6,
  kBufEdgeInfo = 7,
  kBufStage = 8,
  kBufMisc = 9,
  kBufFlags = 10,
  kBufTrace = 11,
  kBufWalkRec = 12,
  kBufWalkFill = 13,
  kBufSubset = 14,  // source list of tgtp_score_sources
  kBufMtEll = 15,   // MT19937 2-way walks: rounds per source
  kBufMtOff = 16,   // ... stream offset per CSR position
  kBufMtDraws = 17, // ... the stream
  kBufPartRecA = 18,   // partitioned walks: record pools of two rounds
  kBufPartRecB = 19,
  kBufPartMeta = 20,   // ... their block tables
  kBufPartCopies = 21, // ... per-workgroup counter copies
  kBufPartTaint = 22,  // ... bouquet fallback bitmap
};

// Returns a device buffer of at least `bytes` in slot `slot` (grow-only).
void* ctx_buffer(tgtp_ctx* ctx, int slot, size_t bytes);

int set_error(tgtp_ctx* ctx, int code, const std::string& msg);

#define TGTP_HIP_CHECK(ctx, expr)                                                   \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess)                                                           \
      return set_error((ctx), TGTP_E_HIP,                                           \
                       std::string(#expr) + ": " + hipGetErrorString(_e));          \
  } while (0)

// ---- kernel launchers (defined in the .hip files) ----------------------------
// Walk counters (saba_walk.hip): clear counters [n] and sums [2]; then
// sums[0] += sum(c), sums[1] += sum((v+1) c[v]).
int launch_walk_counters_clear(tgtp_ctx* ctx, uint32_t* d_counters, uint64_t n,
                               unsigned long long* d_sums);
int launch_walk_counter_sums(tgtp_ctx* ctx, const uint32_t* d_counters, uint64_t n,
                             unsigned long long* d_sums);
// MT19937 comparator walks (mt_walk.hip); variant 0 regular, 1 simd.
int launch_mt_walks(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin, uint32_t src_end,
                    uint32_t depth, uint32_t n_walks, int32_t variant, uint32_t* d_counters,
                    unsigned long long* d_sums, float* walk_ms, bool may_alloc);
// The first `count` outputs of std::mt19937(seed), in order (mt_walk.hip).
int launch_mt_stream(tgtp_ctx* ctx, uint32_t seed, uint64_t count, uint32_t* d_out);
// Partitioned walk engine (saba_part.hip) for large graphs: plans the
// partitions (cached on the graph) and reports whether a pass of these
// parameters takes it; then runs a pass, adding into cleared counters.
// Packed walk layout walk_adj = {id | row start | degree} (saba_walk.hip), built
// once per graph when the fields fit 64 bits; *packed reports whether it did.
int ensure_walk_adj(tgtp_ctx* ctx, tgtp_dgraph* dg, bool* packed);
int saba_part_prepare(tgtp_ctx* ctx, tgtp_dgraph* dg, uint64_t n_src, uint32_t depth,
                      uint32_t n_walks, bool* use);
int launch_saba_walks_part(tgtp_ctx* ctx, tgtp_dgraph* dg, uint32_t src_begin, uint64_t n_src,
                           uint32_t depth, uint32_t n_walks, uint32_t* d_counters);
int launch_saba_walks(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin,
                      uint32_t src_end, uint32_t depth, uint32_t n_walks,
                      uint32_t* d_counters, unsigned long long* d_sums,
                      float* walk_ms);
