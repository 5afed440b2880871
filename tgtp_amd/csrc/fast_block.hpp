// fast_block.hpp -- launch interface of the blocked FAST_FP32 / FAST_BF16
// per-source pipeline (fast_block.hip).
//
// Sources are processed in blocks of 64 (one block = up to 64 sources of one
// graph).  A block's diffusion state is a dense n x 64 float matrix
// (vertex-major: the 64 sources' values of one vertex are one 256-byte row),
// and one diffusion round of all 64 sources is one sparse-times-dense
// product with the graph's random-walk matrix.  Blocks are grouped into
// subsets whose state fits the plan's buffers; a subset runs
//   zero + init, sparse first round, (R - 1) x (round, stop), transpose, post
// where R bounds the rounds any source of the subset can take (the largest
// tau of its rows), the stop kernel applies the reference's cost rule per
// source, and post does the top-gamma bound, chi, walk counts and 2-way walks
// per source on its transposed (contiguous) vector.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "common.hpp"

constexpr uint32_t kBlkCols = 64;  // sources per block (one per lane)
constexpr uint32_t kBlkRows = 64;  // vertices per round tile
constexpr uint32_t kPostLds = 12288;  // post kernel: vectors up to this size are staged in LDS

struct BlockInfo {
  uint32_t voff;   // first vertex of the block's graph in the union
  uint32_t n;      // vertices of that graph
  uint32_t src0;   // first source of the block (graph-local id)
  uint32_t cnt;    // sources in the block (<= 64)
  uint32_t gi;     // graph index
  uint32_t tile0;  // tile units of the subset before this block: tiles
                   // (tile_mode 0) or groups of 8 consecutive tiles (mode 1)
  uint64_t xoff;   // float offset of the block's n x 64 state in the subset buffers
  uint32_t hub0;     // hub rows of the subset's blocks before this block
  uint32_t hub_deg;  // rows above this degree are the graph's hub rows (0: none)
};

struct BlockArgs {
  const uint2* rowd;
  const uint32_t* adj;
  const float* degf;
  const int32_t* taus;
  double* preds_d;     // [m] FP64 row sums
  uint64_t* nr_pos;    // [m]
  uint32_t* status;    // bit 0: walk count beyond 2^36; bit 1: round bound exceeded
  const double* g_eps;
  const double* g_logmd;
  const int32_t* g_gamma;
  const BlockInfo* blocks;
  float* xa;           // subset state buffers
  float* xb;
  // per source column, indexed block * 64 + lane
  uint8_t* active;
  int32_t* ell;
  float* gmax;
  float* gmin;
  uint32_t* qsize;
  uint32_t* st_mx;     // round statistics (float bits), reset by the stop kernel
  uint32_t* st_mn;
  uint32_t* st_q;
  unsigned long long* st_dsum;
  uint32_t* last;      // per block: last round any of its sources took
  uint32_t* next_src;  // post kernel work counter
  uint32_t tile_mode;  // 0: per-XCD runs of tiles; 1: per-block groups of 8 tiles
  uint32_t light;      // degree up to which one thread scans a vertex's row
  uint32_t small_bq;   // walk positions with at most this many bouquets: one 8-lane group
  uint32_t round_form; // 0: one row per wave; 1 / 2: four rows per wave, 8 / 16 gathers per step
  uint32_t adj_words;  // CSR entries (the four-rows form needs m < 2^30)
  // rows of degree above a block's hub_deg are summed by hub workgroups;
  // hubs[hub_off[g] ..] lists graph g's, as union vertex ids
  const uint32_t* hubs;
  const uint32_t* hub_off;
  int32_t* tr_ell;
  float* tr_emax;
  float* tr_gmax;
  float* tr_gmin;
  uint32_t* tr_qsize;
  float* tr_x;
};

struct BlockSubset {
  uint32_t b0, b1;     // blocks [b0, b1)
  uint32_t tiles;      // tile units (see BlockInfo::tile0)
  uint32_t rounds;     // rounds to launch (>= every source's stopping round)
  uint32_t round_tiles0;  // index of this subset's per-round tile and hub counts
  uint64_t floats;     // state floats per buffer
  uint32_t post_stage;   // floats of LDS per post workgroup: the largest vector
                         // of the subset up to kPostLds (0: none staged)
  uint32_t post_per_cu;  // post workgroups resident per CU with that LDS
};

// Optional per-kernel timing of launch_block_subset: an event pair on the
// launch's stream around every round, stop and post launch (diagnostic runs
// only; a captured run passes none).
struct LaunchTimer {
  enum Kind { kRound = 0, kStop = 1, kPost = 2, kKinds = 3 };
  std::vector<hipEvent_t> ev;  // pairs: start, end
  std::vector<int> kind;       // one per pair
  int begin(hipStream_t s, int k);  // records a start event, returns the pair index (-1 on failure)
  void end(hipStream_t s, int pair);
  ~LaunchTimer();
};

// Enqueues the whole per-source phase for one subset on `stream`.
// round_tiles[r]: the tile units round r covers (the subset's blocks are
// ordered by decreasing round bound, so they are a prefix).
// round_hubs[r]: the hub rows round r covers (a prefix, as the tiles).
// post_slots: post kernel workgroups (resident ones; they loop over sources).
int launch_block_subset(tgtp_ctx* ctx, hipStream_t stream, const BlockArgs& A,
                        const BlockSubset& S, const uint32_t* round_tiles,
                        const uint32_t* round_hubs, uint32_t post_slots,
                        LaunchTimer* timer = nullptr);
// Resident workgroups per CU of the post kernel launched with `stage`
// floats of LDS vector buffer.
int block_post_per_cu(uint32_t stage);
