// fast_block.hip -- blocked FAST_FP32 / FAST_BF16 per-source pipeline.
//
// Same algorithm as AESC::tgtp (aesc.h:444-551) with the round structure
// made synchronous:
//   * diffusion round l -> l+1 (the synchronous form of aesc.h:219-239):
//     pvec'[u] = (sum_{w in adj(u)} pvec[w]) / deg(u), so pvec_l = P^l e_s
//     (P = D^-1 A), the series TGT+ truncates.  The frontier of round l is
//     the support of pvec_l.  64 sources advance together: their vectors are
//     the columns of an n x 64 block, a round is one sparse x dense product,
//     and each neighbour read is one 256-byte row (one float per source),
//     instead of one random 4-byte read per source.  Each sum runs in
//     adjacency order in one lane: deterministic, and independent of which
//     sources share a block.
//   * stopping rule per source as the reference (RWRATIO 25, aesc.h:493-503),
//     with the round's max / min / size / degree sum of the frontier; a
//     stopped source's column is carried unchanged through later rounds.
//   * top-gamma bound (ties at the gamma-th value keep every tied vertex),
//     chi and walk counts (aesc.h:392-442, 505-521), then the 2-way SABA
//     walks with the reference's grid / hash / fallback per bouquet; each
//     bouquet's 8 lane sums are added in FP64, bouquets in a fixed order;
//     per-edge row sums (preds) are FP64 throughout.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "common.hpp"
#include "fast_block.hpp"
#include "wave.hpp"

using namespace tgtp_dev;

namespace {

constexpr int kPB = 256;               // post kernel: threads per source
constexpr int kPW = kPB / 64;
constexpr uint32_t kTopK = 16;         // single-pass top-gamma for gamma <= 16
constexpr uint32_t kPassU = 8;         // loads per thread in flight in passes over a vector
constexpr double kMaxWalksFast = 68719476736.0;
constexpr uint32_t kInfBits = 0x7f800000u;

// Last block in [b0, b1) whose first tile unit is <= t.
__device__ __forceinline__ uint32_t find_block(const BlockInfo* B, uint32_t b0, uint32_t b1,
                                               uint32_t t) {
  uint32_t lo = b0, hi = b1;
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (B[mid].tile0 <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Tile of this workgroup (returns false for a workgroup past the last tile;
// workgroups are dealt to the 8 XCDs round-robin, blockIdx.x % 8).
//  * tile_mode 0 (blocks of up to 65,536 vertices): XCD x takes the x-th
//    contiguous eighth of the subset's tiles, so each XCD works through
//    whole blocks of its own and its L2 holds their hot rows;
//  * tile_mode 1 (larger blocks, whose n x 64 states would not fit the
//    Infinity Cache eight at a time): tile groups -- 8 consecutive tiles of
//    one block, one per XCD -- run in block order, so all XCDs share one
//    block's state.
// tile0 of a block counts tiles (mode 0) or tile groups (mode 1) before it.
// vb / vgrid: the workgroup's index among the launch's tile workgroups and
// their number (a multiple of 8; workgroups before them may do other work,
// in multiples of 8, so that vb % 8 still names one XCD).
__device__ __forceinline__ bool tile_of(const BlockInfo* blocks, uint32_t b0, uint32_t b1,
                                        uint32_t n_units, uint32_t mode, uint32_t* bo,
                                        uint32_t* row0, uint32_t vb, uint32_t vgrid) {
  uint32_t unit, t;
  if (mode == 0) {
    const uint32_t per = vgrid >> 3;
    unit = (vb & 7u) * per + (vb >> 3);
    if (unit >= n_units) return false;
  } else {
    unit = vb >> 3;
    if (unit >= n_units) return false;
  }
  const uint32_t b = find_block(blocks, b0, b1, unit);
  const uint32_t tiles = (blocks[b].n + kBlkRows - 1) / kBlkRows;
  t = mode == 0 ? unit - blocks[b].tile0 : 8u * (unit - blocks[b].tile0) + (vb & 7u);
  if (t >= tiles) return false;
  *bo = b;
  *row0 = t * kBlkRows;
  return true;
}

__global__ void blk_zero_kernel(float* __restrict__ x, uint64_t n) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) x[i] = 0.0f;
}

// One wave per source column: column state, e_s, preds row = 1/d_s.
__global__ void __launch_bounds__(256) blk_init_kernel(BlockArgs A, uint32_t b0, uint32_t b1) {
  const uint32_t c = blockIdx.x * 4u + wave_id();
  const uint32_t b = b0 + c / kBlkCols, j = c % kBlkCols;
  if (b >= b1) return;
  const BlockInfo B = A.blocks[b];
  const uint32_t col = b * kBlkCols + j;
  if (lane_id() == 0) {
    A.st_mx[col] = 0u;
    A.st_mn[col] = kInfBits;
    A.st_q[col] = 0u;
    A.st_dsum[col] = 0ull;
    A.ell[col] = 0;
    A.gmax[col] = 0.0f;
    A.gmin[col] = 0.0f;
    A.qsize[col] = 0u;
    A.active[col] = j < B.cnt ? 1 : 0;
    if (j == 0) A.last[b] = 0u;
  }
  if (j >= B.cnt) return;
  const uint32_t s = B.src0 + j;
  const uint2 r = A.rowd[B.voff + s];
  const double dsd = static_cast<double>(r.y);
  float* X1 = A.xb + B.xoff + j;
  for (uint32_t p = r.x + lane_id(); p < r.x + r.y; p += 64) {
    A.preds_d[p] = 1.0 / dsd;
    // round 1 from e_s touches only N(s): count s's occurrences in each
    // neighbour's row (= u's occurrences in s's row; integers, exact in any
    // order), divided by deg(u) in blk_first_kernel
    atomicAdd(&X1[static_cast<uint64_t>(A.adj[p] - B.voff) * kBlkCols], 1.0f);
  }
}

// Round 1: pvec_1[u] = count(s in adj(u)) / deg(u) on N(s) (the value the
// dense round computes: a sum of ones in adjacency order, then the same
// division), with the round's frontier statistics.  The first occurrence of
// u in s's sorted row divides; duplicates are skipped.
__global__ void __launch_bounds__(256) blk_first_kernel(BlockArgs A, uint32_t b0, uint32_t b1) {
  const uint32_t c = blockIdx.x * 4u + wave_id();
  const uint32_t b = b0 + c / kBlkCols, j = c % kBlkCols;
  if (b >= b1) return;
  const BlockInfo B = A.blocks[b];
  if (j >= B.cnt) return;
  const uint32_t col = b * kBlkCols + j;
  const uint32_t s = B.src0 + j;
  const uint2 r = A.rowd[B.voff + s];
  float* X1 = A.xb + B.xoff + j;
  float mx = 0.0f, mn = INFINITY;
  uint32_t q = 0;
  unsigned long long ds = 0;
  for (uint32_t p = r.x + lane_id(); p < r.x + r.y; p += 64) {
    const uint32_t ug = A.adj[p];
    if (p > r.x && A.adj[p - 1] == ug) continue;
    float* x = &X1[static_cast<uint64_t>(ug - B.voff) * kBlkCols];
    const float val = *x / A.degf[ug];
    *x = val;
    mx = fmaxf(mx, val);
    mn = fminf(mn, val);
    ++q;
    ds += A.rowd[ug].y;
  }
  mx = wave_max(mx);
  mn = wave_min(mn);
  q = wave_sum(q);
  ds = wave_sum(ds);
  if (lane_id() == 0) {
    A.st_mx[col] = __float_as_uint(mx);
    A.st_mn[col] = __float_as_uint(mn);
    A.st_q[col] = q;
    A.st_dsum[col] = ds;
  }
}

// One round for every tile (64 vertices of one block) of the subset: lane j
// computes source j's new value at each vertex of the tile.  One workgroup
// per tile (a persistent loop over tiles measured slower: the hardware
// overlaps a finishing workgroup with the next one, a loop with a barrier
// per tile does not).
__global__ void __launch_bounds__(256)
blk_round_kernel(BlockArgs A, uint32_t b0, uint32_t b1, uint32_t n_units,
                 const float* __restrict__ cur, float* __restrict__ nxt) {
  __shared__ float s_mx[4][kBlkCols];
  __shared__ float s_mn[4][kBlkCols];
  __shared__ uint32_t s_q[4][kBlkCols];
  __shared__ unsigned long long s_ds[4][kBlkCols];
  uint32_t b, row0;
  if (!tile_of(A.blocks, b0, b1, n_units, A.tile_mode, &b, &row0, blockIdx.x, gridDim.x))
    return;  // uniform per wg
  const BlockInfo B = A.blocks[b];
  const uint32_t j = lane_id(), wv = wave_id();
  const uint32_t col = b * kBlkCols + j;
  const bool act = A.active[col] != 0;
  // the same flags in every wave: a block whose sources all stopped is
  // skipped by the whole workgroup
  if (__ballot(act) == 0ull) return;
  float mx = 0.0f, mn = INFINITY;
  uint32_t q = 0;
  unsigned long long ds = 0;
  const uint32_t rend = min(row0 + kBlkRows, B.n);
  const float* Xc = cur + B.xoff;
  float* Xn = nxt + B.xoff;
  // row descriptors and neighbour ids are the same for the whole wave:
  // made uniform, they are scalar loads and each neighbour costs one
  // 256-byte vector load of its state row
  const uint32_t wu = __builtin_amdgcn_readfirstlane(wv);
  // (Two rows per wave with interleaved gathers measured 8-20% slower.)
  for (uint32_t u = row0 + wu; u < rend; u += 4) {
    const uint64_t ro = static_cast<uint64_t>(u) * kBlkCols + j;
    // rows above the block's hub degree belong to blk_hub_kernel
    if (B.hub_deg != 0 && A.rowd[B.voff + u].y > B.hub_deg) continue;  // uniform per wave
    if (act) {
      const uint2 r = A.rowd[B.voff + u];
      const uint32_t start = __builtin_amdgcn_readfirstlane(r.x);
      const uint32_t deg = __builtin_amdgcn_readfirstlane(r.y);
      const uint32_t* nb = A.adj + start;
      const float* xj = Xc + j;
      const uint32_t voff = B.voff;
      float sum = 0.0f;
      uint32_t k = 0;
      // loads of a chunk are issued together, the adds stay in adjacency order
      for (; k + 16 <= deg; k += 16) {
        float v[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) v[c] = xj[static_cast<uint64_t>(nb[k + c] - voff) * kBlkCols];
#pragma unroll
        for (int c = 0; c < 16; ++c) sum += v[c];
      }
      for (; k + 4 <= deg; k += 4) {
        float v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = xj[static_cast<uint64_t>(nb[k + c] - voff) * kBlkCols];
#pragma unroll
        for (int c = 0; c < 4; ++c) sum += v[c];
      }
      for (; k < deg; ++k) sum += xj[static_cast<uint64_t>(nb[k] - voff) * kBlkCols];
      const float val = sum / A.degf[B.voff + u];
      Xn[ro] = val;
      if (val > 0.0f) {
        mx = fmaxf(mx, val);
        mn = fminf(mn, val);
        ++q;
        ds += r.y;
      }
    } else {
      Xn[ro] = Xc[ro];
    }
  }
  s_mx[wv][j] = mx;
  s_mn[wv][j] = mn;
  s_q[wv][j] = q;
  s_ds[wv][j] = ds;
  __syncthreads();
  if (wv == 0) {
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      mx = fmaxf(mx, s_mx[w][j]);
      mn = fminf(mn, s_mn[w][j]);
      q += s_q[w][j];
      ds += s_ds[w][j];
    }
    if (act && q) {
      atomicMax(&A.st_mx[col], __float_as_uint(mx));  // non-negative floats order as bits
      atomicMin(&A.st_mn[col], __float_as_uint(mn));
      atomicAdd(&A.st_q[col], q);
      atomicAdd(&A.st_dsum[col], ds);
    }
  }
}

constexpr uint32_t kOutOfRange = 0xfffffff0u;  // buffer offset past any state
constexpr int kRsrcWord3 = 0x00020000;         // gfx9 raw buffer: 32-bit data, no swizzle

// kC neighbours k .. k + kC - 1 of the quarter's row added to this lane's
// four sums, with lane ql of the quarter holding the id of neighbour k + ql
// in `my`: the ids are read across the quarter, then all kC 16-byte gathers
// are issued before the adds.  Slots past the row's degree (and lanes whose
// four sources have all stopped) read out of range: no memory access, and
// the zeros they return leave the non-negative sums unchanged.
template <int kC>
__device__ __forceinline__ void gather_ids(__amdgpu_buffer_rsrc_t xrs, uint32_t my, uint32_t deg,
                                           bool any, uint32_t l, uint32_t k, float4& sum) {
  const uint32_t ql = l & 15u;
  uint32_t off[kC];
#pragma unroll
  for (int c = 0; c < kC; ++c) {
    const uint32_t id = __shfl(my, static_cast<int>((l & 48u) | static_cast<uint32_t>(c)));
    off[c] = any && k + c < deg ? id * (kBlkCols * 4u) + ql * 16u : kOutOfRange;
  }
  float4 v[kC];
#pragma unroll
  for (int c = 0; c < kC; ++c)
    v[c] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(xrs, off[c], 0, 0));
#pragma unroll
  for (int c = 0; c < kC; ++c) {
    sum.x += v[c].x;
    sum.y += v[c].y;
    sum.z += v[c].z;
    sum.w += v[c].w;
  }
}

// Lane ql's neighbour id k + ql of a row (block-local; anything past the
// row's first 16 from k, or past its degree, reads out of range).
__device__ __forceinline__ uint32_t row_id(__amdgpu_buffer_rsrc_t ars, uint2 r, uint32_t voff,
                                           uint32_t ql, uint32_t k) {
  const uint32_t off = k + ql < r.y ? (r.x + k + ql) * 4u : kOutOfRange;
  return __builtin_amdgcn_raw_buffer_load_b32(ars, off, 0, 0) - voff;
}

// Last block in [b0, b1) whose first hub row is <= t.
__device__ __forceinline__ uint32_t find_hub_block(const BlockInfo* B, uint32_t b0, uint32_t b1,
                                                   uint32_t t) {
  uint32_t lo = b0, hi = b1;
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (B[mid].hub0 <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// One round of one hub row (degree above the block's hub_deg) of one block,
// by a whole workgroup: the row's neighbours in 16 contiguous segments, one
// per quarter-wave (lane ql holding sources 4 ql .. 4 ql + 3), each summed in
// adjacency order; the 16 segment sums are then added in segment order.  A
// long row thus costs ~deg / 256 dependent gather steps instead of deg / 16
// in one quarter-wave.  `unit` indexes the hub rows of the active blocks.
template <int kChunk>
__device__ void hub_row(const BlockArgs& A, uint32_t b0, uint32_t b1, uint32_t unit,
                        const float* __restrict__ cur, float* __restrict__ nxt,
                        float4 (*part)[16]) {
  constexpr uint32_t kSeg = 16;
  const uint32_t b = find_hub_block(A.blocks, b0, b1, unit);
  const BlockInfo B = A.blocks[b];
  const uint32_t voff = B.voff;
  const uint32_t v = A.hubs[A.hub_off[B.gi] + (unit - B.hub0)] - voff;
  const uint32_t l = lane_id(), wv = wave_id();
  const uint32_t qd = l >> 4, ql = l & 15u;
  const uchar4 a4 = *reinterpret_cast<const uchar4*>(A.active + b * kBlkCols + ql * 4u);
  const bool any = a4.x || a4.y || a4.z || a4.w;
  if (__ballot(any) == 0ull) return;  // same flags in every wave
  const uint2 r = A.rowd[voff + v];
  const uint32_t len = (r.y + kSeg - 1) / kSeg;
  const uint32_t seg = wv * 4u + qd;
  const uint32_t k0 = min(r.y, seg * len), k1 = min(r.y, k0 + len);
  const __amdgpu_buffer_rsrc_t xrs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(cur + B.xoff), 0, static_cast<int>(B.n * (kBlkCols * 4u)), kRsrcWord3);
  // the row's own ids as the buffer (no 32-bit offset limit on the union's
  // CSR; the host gives hub rows only to graphs whose rows all have fewer
  // than 2^29 entries): ids past the segment read out of range
  const __amdgpu_buffer_rsrc_t ars = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint32_t*>(A.adj + r.x), 0, static_cast<int>(r.y * 4u), kRsrcWord3);
  const uint2 seg_row = make_uint2(0u, k1);
  float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t my = row_id(ars, seg_row, voff, ql, k0);
  for (uint32_t k = k0; k < k1; k += kChunk) {
    const uint32_t my_n = row_id(ars, seg_row, voff, ql, k + kChunk);
    gather_ids<kChunk>(xrs, my, k1, any, l, k, sum);
    my = my_n;
  }
  part[seg][ql] = sum;
  __syncthreads();
  if (wv != 0) return;
  const uint32_t j = l, col = b * kBlkCols + j;
  float t = 0.0f;
#pragma unroll
  for (uint32_t q = 0; q < kSeg; ++q) {
    const float4 p4 = part[q][j >> 2];
    t += (j & 3u) == 0 ? p4.x : ((j & 3u) == 1 ? p4.y : ((j & 3u) == 2 ? p4.z : p4.w));
  }
  const uint64_t ro = static_cast<uint64_t>(v) * kBlkCols + j;
  const bool act = A.active[col] != 0;
  const float val = act ? t / static_cast<float>(r.y) : cur[B.xoff + ro];
  nxt[B.xoff + ro] = val;
  if (act && val > 0.0f) {
    atomicMax(&A.st_mx[col], __float_as_uint(val));
    atomicMin(&A.st_mn[col], __float_as_uint(val));
    atomicAdd(&A.st_q[col], 1u);
    atomicAdd(&A.st_dsum[col], static_cast<unsigned long long>(r.y));
  }
}

// The same round with four rows per wave: each 16-lane quarter takes one
// row and each lane four sources (16-byte loads), so one load instruction
// gathers a neighbour row for four vertices at once and a wave keeps up to
// 4 x kChunk gathers in flight on low-degree rows, where the one-row form
// has only deg of them.  Each source's sum runs in adjacency order exactly
// as in blk_round_kernel, so both forms give the same bits.
template <int kChunk, int kWaves>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWaves)))
blk_round4_kernel(BlockArgs A, uint32_t b0, uint32_t b1, uint32_t n_units, uint32_t n_hubs,
                  const float* __restrict__ cur, float* __restrict__ nxt) {
  constexpr int kQuarters = 16;  // 4 waves x 4 quarters
  __shared__ float s_mx[kQuarters][kBlkCols];
  __shared__ float s_mn[kQuarters][kBlkCols];
  __shared__ uint32_t s_q[kQuarters][kBlkCols];
  __shared__ unsigned long long s_ds[kQuarters][kBlkCols];
  __shared__ float4 s_part[16][16];
  // the first workgroups (n_hubs rounded up to 8) take the hub rows, whose
  // long chains then overlap the tiles (16-gather form; the 8-gather form
  // has n_hubs = 0 and blk_hub_kernel after it: the hub path would spill
  // there)
  const uint32_t hub_wgs = (n_hubs + 7u) & ~7u;
  if constexpr (kChunk >= 16) {
    if (blockIdx.x < hub_wgs) {
      if (blockIdx.x < n_hubs) hub_row<16>(A, b0, b1, blockIdx.x, cur, nxt, s_part);
      return;
    }
  }
  uint32_t b, row0;
  if (!tile_of(A.blocks, b0, b1, n_units, A.tile_mode, &b, &row0, blockIdx.x - hub_wgs,
               gridDim.x - hub_wgs))
    return;  // uniform per wg
  const BlockInfo B = A.blocks[b];
  const uint32_t l = lane_id(), wv = wave_id();
  const uint32_t qd = l >> 4, ql = l & 15u, c0 = ql * 4u;
  const uchar4 a4 = *reinterpret_cast<const uchar4*>(A.active + b * kBlkCols + c0);
  const bool act[4] = {a4.x != 0, a4.y != 0, a4.z != 0, a4.w != 0};
  const bool any = act[0] || act[1] || act[2] || act[3];
  const bool all = act[0] && act[1] && act[2] && act[3];
  if (__ballot(any) == 0ull) return;  // same flags in every wave
  float mx[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  float mn[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  uint32_t q[4] = {0, 0, 0, 0};
  unsigned long long ds[4] = {0, 0, 0, 0};
  const uint32_t rend = min(row0 + kBlkRows, B.n);
  float* Xn = nxt + B.xoff + c0;
  const uint32_t voff = B.voff;
  // the block's state as a buffer resource (n x 256 bytes; the host keeps
  // this form to n < 2^24): a gather is a 16-byte buffer load at lane offset
  // id * 256 + 16 * ql.  The adjacency likewise (m < 2^30).
  const __amdgpu_buffer_rsrc_t xrs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(cur + B.xoff), 0, static_cast<int>(B.n * (kBlkCols * 4u)), kRsrcWord3);
  const __amdgpu_buffer_rsrc_t ars = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint32_t*>(A.adj), 0, static_cast<int>(A.adj_words * 4u), kRsrcWord3);
  // quarter qd of wave wv takes rows row0 + 16 g + 4 wv + qd, g = 0..3: a
  // wave's four rows are consecutive (similar degrees) and its stores one
  // 1 KB run.  Software pipeline over g: the four row descriptors load
  // first, and the next group's first 16 neighbour ids load while this
  // group's gathers are in flight.
  constexpr uint32_t kGroups = kBlkRows / 16u;
  const uint32_t base0 = row0 + 4u * __builtin_amdgcn_readfirstlane(wv);
  uint2 rr[kGroups];
  {
    // buffer loads (rows past the tile read out of range as {0, 0}): no
    // branches, so all four are in flight at once
    const __amdgpu_buffer_rsrc_t rrs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint2*>(A.rowd + voff), 0, static_cast<int>(B.n * 8u), kRsrcWord3);
#pragma unroll
    for (uint32_t g = 0; g < kGroups; ++g) {
      const uint32_t u = base0 + 16u * g + qd;
      rr[g] = __builtin_bit_cast(
          uint2, __builtin_amdgcn_raw_buffer_load_b64(rrs, u < rend ? u * 8u : kOutOfRange, 0, 0));
    }
  }
  uint32_t my_next = row_id(ars, rr[0], voff, ql, 0);
#pragma unroll
  for (uint32_t g = 0; g < kGroups; ++g) {
    const uint32_t base = base0 + 16u * g;
    if (base >= rend) break;  // uniform
    const uint32_t u = base + qd;
    const uint2 r = rr[g];
    // rows above the block's hub degree belong to the hub workgroups (hub_row)
    const bool rv = u < rend && (B.hub_deg == 0 || r.y <= B.hub_deg);
    const uint32_t my0 = my_next;
    if (g + 1 < kGroups) my_next = row_id(ars, rr[g + 1], voff, ql, 0);
    // the row's degree is the same in all 16 lanes of the quarter, whatever
    // their sources' flags: every lane fetches its share of the ids
    const uint32_t deg = rv ? r.y : 0u;
    const uint32_t dmax = max(max(__builtin_amdgcn_readlane(deg, 0), __builtin_amdgcn_readlane(deg, 16)),
                              max(__builtin_amdgcn_readlane(deg, 32), __builtin_amdgcn_readlane(deg, 48)));
    const uint64_t ro = static_cast<uint64_t>(u) * kBlkCols;
    // stopped sources keep their value: read with the gathers (out of range,
    // so no access, when all four of the lane's sources are active)
    const float4 old = __builtin_bit_cast(
        float4, __builtin_amdgcn_raw_buffer_load_b128(
                    xrs, rv && !all ? static_cast<uint32_t>(ro) * 4u + ql * 16u : kOutOfRange, 0, 0));
    float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // steps of kChunk gathers while a row of the wave has more than half a
    // chunk left, then of 8 and 4 (fewer idle slots on short rows); the
    // first step's ids are the prefetched ones
    // (each step's ids load during the previous step's gathers)
    auto step_of = [&](uint32_t k) -> uint32_t {
      const uint32_t rem = dmax - k;
      return kChunk >= 16 && rem > 8 ? 16u : (rem > 4 ? 8u : 4u);
    };
    uint32_t k = 0, my = my0;
    while (k < dmax) {
      const uint32_t st = step_of(k);
      const uint32_t kn = k + st;
      // (unconditional: past the degree it reads out of range)
      const uint32_t my_n = row_id(ars, r, voff, ql, kn);
      if (st == 16u)
        gather_ids<16>(xrs, my, deg, any, l, k, sum);
      else if (st == 8u)
        gather_ids<8>(xrs, my, deg, any, l, k, sum);
      else
        gather_ids<4>(xrs, my, deg, any, l, k, sum);
      k = kn;
      my = my_n;
    }
    if (rv) {
      // the degree as a float, as the graph's degf table holds it
      const float d = static_cast<float>(deg);
      float4 val = make_float4(sum.x / d, sum.y / d, sum.z / d, sum.w / d);
      if (!act[0]) val.x = old.x;
      if (!act[1]) val.y = old.y;
      if (!act[2]) val.z = old.z;
      if (!act[3]) val.w = old.w;
      *reinterpret_cast<float4*>(Xn + ro) = val;
      const float vv[4] = {val.x, val.y, val.z, val.w};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (act[i] && vv[i] > 0.0f) {
          mx[i] = fmaxf(mx[i], vv[i]);
          mn[i] = fminf(mn[i], vv[i]);
          ++q[i];
          ds[i] += r.y;
        }
      }
    }
  }
  const uint32_t slot = wv * 4u + qd;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    s_mx[slot][c0 + i] = mx[i];
    s_mn[slot][c0 + i] = mn[i];
    s_q[slot][c0 + i] = q[i];
    s_ds[slot][c0 + i] = ds[i];
  }
  __syncthreads();
  if (wv == 0) {
    const uint32_t j = l;
    float fx = s_mx[0][j], fn = s_mn[0][j];
    uint32_t fq = s_q[0][j];
    unsigned long long fd = s_ds[0][j];
#pragma unroll
    for (int w = 1; w < kQuarters; ++w) {
      fx = fmaxf(fx, s_mx[w][j]);
      fn = fminf(fn, s_mn[w][j]);
      fq += s_q[w][j];
      fd += s_ds[w][j];
    }
    const uint32_t col = b * kBlkCols + j;
    if (fq && A.active[col]) {
      atomicMax(&A.st_mx[col], __float_as_uint(fx));
      atomicMin(&A.st_mn[col], __float_as_uint(fn));
      atomicAdd(&A.st_q[col], fq);
      atomicAdd(&A.st_dsum[col], fd);
    }
  }
}

// The hub rows of a round on their own (with the one-row and 8-gather forms).
__global__ void __launch_bounds__(256)
blk_hub_kernel(BlockArgs A, uint32_t b0, uint32_t b1, uint32_t n_units,
               const float* __restrict__ cur, float* __restrict__ nxt) {
  __shared__ float4 part[16][16];
  if (blockIdx.x < n_units) hub_row<16>(A, b0, b1, blockIdx.x, cur, nxt, part);
}

// After a round: per active source, the round's row update of preds and the
// reference's cost rule (aesc.h:493-503).  One wave per source column.
__global__ void __launch_bounds__(256)
blk_stop_kernel(BlockArgs A, uint32_t b0, uint32_t b1, const float* __restrict__ xn) {
  const uint32_t c = blockIdx.x * 4u + wave_id();
  const uint32_t b = b0 + c / kBlkCols, j = c % kBlkCols;
  if (b >= b1) return;
  const BlockInfo B = A.blocks[b];
  const uint32_t col = b * kBlkCols + j;
  if (j >= B.cnt || A.active[col] == 0) return;
  const float mx = __uint_as_float(A.st_mx[col]);
  const float mn = __uint_as_float(A.st_mn[col]);
  const uint32_t q = A.st_q[col];
  const unsigned long long dsum = A.st_dsum[col];
  const int ell = A.ell[col] + 1;
  const float gmax = mx;
  const float gmin = q < B.n ? 0.0f : mn;
  const uint32_t s = B.src0 + j;
  const uint2 r = A.rowd[B.voff + s];
  const double dsd = static_cast<double>(r.y);
  const double eps = A.g_eps[B.gi], logmd = A.g_logmd[B.gi];
  const double dse_d = dsd * eps;
  const double dse2 = dse_d * dse_d;
  const float* X = xn + B.xoff;
  const double ps = static_cast<double>(X[static_cast<uint64_t>(s) * kBlkCols + j]);
  double rw = 0.0;
  auto walk_estimate = [&](int rest) {
    if (rest > 0) {
      const double chi = 2.0 * rest * (static_cast<double>(gmax) - gmin);
      rw += fmax(ceil(8 * (chi * chi) * logmd / dse2), 1.0);
    }
  };
  // a lane's positions four at a time (a hub source's row is otherwise a
  // long chain of dependent loads); the lane's sum keeps position order
  constexpr uint32_t kU = 4;
  const uint32_t pend = r.x + r.y;
  uint32_t p = r.x + lane_id();
  for (; p + 64u * (kU - 1) < pend; p += 64u * kU) {
    uint32_t v[kU];
    int tr[kU];
    double pd[kU];
    float xv[kU];
#pragma unroll
    for (uint32_t i = 0; i < kU; ++i) {
      v[i] = A.adj[p + 64u * i] - B.voff;
      tr[i] = A.taus[p + 64u * i];
      pd[i] = A.preds_d[p + 64u * i];
    }
#pragma unroll
    for (uint32_t i = 0; i < kU; ++i) xv[i] = X[static_cast<uint64_t>(v[i]) * kBlkCols + j];
#pragma unroll
    for (uint32_t i = 0; i < kU; ++i) {
      A.preds_d[p + 64u * i] = pd[i] + (ps - static_cast<double>(xv[i])) / dsd;
      walk_estimate(tr[i] - ell);
    }
  }
  for (; p < pend; p += 64) {
    const uint32_t v = A.adj[p] - B.voff;
    A.preds_d[p] += (ps - static_cast<double>(X[static_cast<uint64_t>(v) * kBlkCols + j])) / dsd;
    walk_estimate(A.taus[p] - ell);
  }
  rw = wave_sum(rw);
  if (lane_id() == 0) {
    A.ell[col] = ell;
    A.gmax[col] = gmax;
    A.gmin[col] = gmin;
    A.qsize[col] = q;
    if (static_cast<double>(dsum) >= 25 * rw) A.active[col] = 0;
    atomicMax(&A.last[b], static_cast<uint32_t>(ell));
    A.st_mx[col] = 0u;
    A.st_mn[col] = kInfBits;
    A.st_q[col] = 0u;
    A.st_dsum[col] = 0ull;
  }
}

// Final block state (the buffer its last round wrote) -> one contiguous
// vector per source in the other buffer.  Also flags a source still active
// after the last launched round (the round bound was wrong).
__global__ void __launch_bounds__(256)
blk_transpose_kernel(BlockArgs A, uint32_t b0, uint32_t b1, uint32_t n_units) {
  __shared__ float tile[kBlkRows][kBlkCols + 1];
  uint32_t b, row0;
  if (!tile_of(A.blocks, b0, b1, n_units, A.tile_mode, &b, &row0, blockIdx.x, gridDim.x))
    return;  // uniform per wg
  const BlockInfo B = A.blocks[b];
  const uint32_t L = A.last[b];
  const float* F = ((L & 1u) ? A.xb : A.xa) + B.xoff;
  float* O = ((L & 1u) ? A.xa : A.xb) + B.xoff;
  const uint32_t lane = lane_id(), wv = wave_id();
  for (uint32_t r = wv; r < kBlkRows; r += 4) {
    const uint32_t u = row0 + r;
    tile[r][lane] = u < B.n ? F[static_cast<uint64_t>(u) * kBlkCols + lane] : 0.0f;
  }
  __syncthreads();
  const uint32_t u = row0 + lane;
  for (uint32_t jj = wv; jj < B.cnt; jj += 4)
    if (u < B.n) O[static_cast<uint64_t>(jj) * B.n + u] = tile[lane][jj];
  if (row0 == 0 && wv == 0 && lane < B.cnt && A.active[b * kBlkCols + lane] != 0)
    atomicOr(A.status, 2u);
}

// ---- post: top-gamma bound, chi, walk counts, 2-way walks -------------------

constexpr uint32_t kCandCap = 1024;  // listed members of the top-gamma set

struct PostShared {
  float rf[kPW];
  double rd[kPW];
  uint32_t ru[kPW];
  uint32_t heavy[kPB];
  uint32_t n_heavy;
  uint32_t k;
  uint32_t cand[kCandCap];
  uint32_t n_cand;
  uint32_t list_all;
  double wacc[kPB / 8];   // per 8-lane group walk sums
};

// One bouquet of the 2-way walk for lane b of an 8-lane group; returns the
// lane's float sum over the bouquet's hops.  The group's 8 lanes run in
// lockstep (same position, bouquet and hop counts).
__device__ __forceinline__ float walk2_bouquet(const BlockArgs& A, const float* pv, uint32_t off,
                                               uint32_t s, uint32_t v, uint32_t j, uint32_t b,
                                               uint32_t step, uint32_t len, uint32_t giw) {
  uint32_t c1 = s, c2 = v;
  uint32_t s1 = (j * 8u + b + 1u) * step;
  uint32_t s2 = s1, st = s1;
  bool fb = false;
  float x = 0.0f;
  for (uint32_t h = 0; h < len; ++h) {
    const uint64_t z = __ballot(!fb && (s1 == 0u || s2 == 0u));
    fb = fb || (((z >> (giw * 8u)) & 0xFFull) != 0ull);
    const uint2 d1 = A.rowd[c1];
    const uint2 d2 = A.rowd[c2];
    uint32_t k1, k2;
    if (!fb) {
      k1 = __umulhi(s1 ^ fmix32(c1 - off), d1.y);
      k2 = __umulhi(s2 ^ fmix32(c2 - off), d2.y);
      s1 *= d1.y;
      s2 *= d2.y;
    } else {
      const uint32_t ra = splitmix32(st);
      const uint32_t rb = splitmix32(st);
      k1 = ra % d1.y;
      k2 = rb % d2.y;
    }
    c1 = A.adj[d1.x + k1];
    c2 = A.adj[d2.x + k2];
    x += pv[c1 - off] - pv[c2 - off];
  }
  return x;
}

// FP64 sum of the 8 lanes of a group, lanes 0..7 in order.
__device__ __forceinline__ double group_sum8(float x, uint32_t grp_base) {
  double s = 0.0;
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) s += static_cast<double>(__shfl(x, static_cast<int>(grp_base + k)));
  return s;
}

// gamma-th largest value of the frontier (pv > 0), counting ties; -inf when
// the frontier is empty.  cand: kPB * kTopK floats of LDS.  For rg <= kTopK
// each thread's kTopK largest values (of the vertices it visits) and their
// vertices are left in top / tix, largest first.
__device__ float top_gamma_value(const float* pv, uint32_t n, uint32_t rg, float* cand,
                                 PostShared& sh, float (&top)[kTopK], uint32_t (&tix)[kTopK]) {
  float thr = INFINITY, vstar = -INFINITY;
  uint32_t above = 0;
  if (rg <= kTopK) {
#pragma unroll
    for (uint32_t i = 0; i < kTopK; ++i) {
      top[i] = -INFINITY;
      tix[i] = 0;
    }
    // kPassU values per thread loaded before any is inserted: on vectors
    // read from global memory the loads overlap (the insertion's branch
    // would otherwise wait for each); the top-16 multiset does not depend
    // on the order of insertion
    for (uint32_t u0 = threadIdx.x; u0 < n; u0 += kPB * kPassU) {
      float xs[kPassU];
#pragma unroll
      for (uint32_t c = 0; c < kPassU; ++c) {
        const uint32_t u = u0 + c * kPB;
        xs[c] = u < n ? pv[u] : 0.0f;
      }
#pragma unroll
      for (uint32_t c = 0; c < kPassU; ++c) {
        const float x = xs[c];
        const uint32_t u = u0 + c * kPB;
        if (x > 0.0f && x > top[kTopK - 1]) {
#pragma unroll
          for (uint32_t i = kTopK - 1; i > 0; --i) {
            const bool up = x > top[i - 1], here = x > top[i];
            tix[i] = up ? tix[i - 1] : (here ? u : tix[i]);
            top[i] = up ? top[i - 1] : (here ? x : top[i]);
          }
          if (x > top[0]) {
            top[0] = x;
            tix[0] = u;
          }
        }
      }
    }
#pragma unroll
    for (uint32_t i = 0; i < kTopK; ++i) cand[threadIdx.x * kTopK + i] = top[i];
    __syncthreads();
    for (uint32_t it = 0; it <= rg; ++it) {
      float m1 = -INFINITY;
      for (uint32_t i = threadIdx.x; i < kPB * kTopK; i += kPB) {
        const float x = cand[i];
        if (x > 0.0f && x < thr) m1 = fmaxf(m1, x);
      }
      m1 = block_reduce<kPW>(m1, sh.rf, OpMaxF{});
      uint32_t c = 0;
      for (uint32_t i = threadIdx.x; i < kPB * kTopK; i += kPB) c += (cand[i] > 0.0f && cand[i] == m1);
      c = block_reduce<kPW>(c, sh.ru, OpSum{});
      vstar = m1;
      if (above + c >= rg || c == 0) break;
      above += c;
      thr = m1;
    }
    return vstar;
  }
  for (uint32_t it = 0; it <= rg; ++it) {  // gamma > 16: threshold passes over pv
    float m1 = -INFINITY;
    for (uint32_t u = threadIdx.x; u < n; u += kPB) {
      const float x = pv[u];
      if (x > 0.0f && x < thr) m1 = fmaxf(m1, x);
    }
    m1 = block_reduce<kPW>(m1, sh.rf, OpMaxF{});
    uint32_t c = 0;
    for (uint32_t u = threadIdx.x; u < n; u += kPB) c += (pv[u] > 0.0f && pv[u] == m1);
    c = block_reduce<kPW>(c, sh.ru, OpSum{});
    vstar = m1;
    if (above + c >= rg || c == 0) break;
    above += c;
    thr = m1;
  }
  return vstar;
}

__device__ void post_source(const BlockArgs& A, const BlockInfo& B, uint32_t col, uint32_t s,
                            const float* pv, float* cand, PostShared& sh) {
  const uint32_t n = B.n, off = B.voff;
  const uint32_t src = off + s;
  const double eps = A.g_eps[B.gi], logmd = A.g_logmd[B.gi];
  const int32_t gamma = A.g_gamma[B.gi];
  const int ell = A.ell[col];
  const float gmax = A.gmax[col], gmin = A.gmin[col];
  const uint32_t Q = A.qsize[col];
  const uint2 rs = A.rowd[src];
  const uint32_t r0 = rs.x, r1 = rs.x + rs.y;
  const double dsd = static_cast<double>(rs.y);
  const double dse_d = dsd * eps;

  // ---- top-gamma bound (aesc.h:414-442)
  const uint32_t rg = min(static_cast<uint32_t>(gamma), Q);
  float top[kTopK];
  uint32_t tix[kTopK];
  const float vstar = top_gamma_value(pv, n, rg, cand, sh, top, tix);
  const float gamma_max = gamma < static_cast<int>(Q) ? fminf(1.0f, vstar) : 0.0f;
  float emax = gmax + gamma_max;
  {
    // the top-gamma set with its ties: listed from the threads' largest
    // values (or by a pass over the vector), then one wave per member scans
    // its row; an overlong tie set takes the chunked scan
    float best = -INFINITY;
    if (threadIdx.x == 0) {
      sh.n_cand = 0;
      sh.list_all = rg > kTopK ? 1u : 0u;
    }
    __syncthreads();
    // a thread's members are among its kTopK largest values unless the
    // smallest of those is itself a member (then ties may hide more): only
    // then is the vector scanned again
    if (rg <= kTopK) {
      if (top[kTopK - 1] > 0.0f && top[kTopK - 1] >= vstar) {
        sh.list_all = 1u;
      } else {
#pragma unroll
        for (uint32_t i = 0; i < kTopK; ++i) {
          if (top[i] > 0.0f && top[i] >= vstar) {
            const uint32_t at = atomicAdd(&sh.n_cand, 1u);
            if (at < kCandCap) sh.cand[at] = tix[i];
          }
        }
      }
    }
    __syncthreads();
    const bool rescan = sh.list_all != 0u;
    if (rescan) {
      __syncthreads();
      if (threadIdx.x == 0) sh.n_cand = 0;
      __syncthreads();
    }
    for (uint32_t u0 = threadIdx.x; rescan && u0 < n; u0 += kPB * kPassU) {
      float xs[kPassU];
#pragma unroll
      for (uint32_t c = 0; c < kPassU; ++c) {
        const uint32_t u = u0 + c * kPB;
        xs[c] = u < n ? pv[u] : 0.0f;
      }
#pragma unroll
      for (uint32_t c = 0; c < kPassU; ++c) {
        if (xs[c] > 0.0f && xs[c] >= vstar) {
          const uint32_t i = atomicAdd(&sh.n_cand, 1u);
          if (i < kCandCap) sh.cand[i] = u0 + c * kPB;
        }
      }
    }
    __syncthreads();
    const uint32_t nc = sh.n_cand;
    if (nc <= kCandCap) {
      for (uint32_t h = wave_id(); h < nc; h += kPW) {
        const uint32_t hu = sh.cand[h];
        const uint2 ru = A.rowd[off + hu];
        const float pu = pv[hu];
        for (uint32_t k = lane_id(); k < ru.y; k += 64) {
          const float pw = pv[A.adj[ru.x + k] - off];
          if (pw > 0.0f && pw >= vstar) best = fmaxf(best, pu + pw);
        }
      }
    }
    for (uint32_t base = 0; nc > kCandCap && base < n; base += kPB) {
      if (threadIdx.x == 0) sh.n_heavy = 0;
      __syncthreads();
      const uint32_t u = base + threadIdx.x;
      if (u < n) {
        const float pu = pv[u];
        if (pu > 0.0f && pu >= vstar) {
          const uint2 ru = A.rowd[off + u];
          if (ru.y <= A.light) {
            for (uint32_t k = 0; k < ru.y; ++k) {
              const float pw = pv[A.adj[ru.x + k] - off];
              if (pw > 0.0f && pw >= vstar) best = fmaxf(best, pu + pw);
            }
          } else {
            sh.heavy[atomicAdd(&sh.n_heavy, 1u)] = u;
          }
        }
      }
      __syncthreads();
      const uint32_t nh = sh.n_heavy;
      for (uint32_t h = wave_id(); h < nh; h += kPW) {
        const uint32_t hu = sh.heavy[h];
        const uint2 ru = A.rowd[off + hu];
        const float pu = pv[hu];
        for (uint32_t k = lane_id(); k < ru.y; k += 64) {
          const float pw = pv[A.adj[ru.x + k] - off];
          if (pw > 0.0f && pw >= vstar) best = fmaxf(best, pu + pw);
        }
      }
      __syncthreads();
    }
    best = block_reduce<kPW>(best, sh.rf, OpMaxF{});
    emax = fmaxf(emax, best);
  }
  if (gamma == 1) emax = 2 * gmax;

  // ---- chi and walk counts, one wave per CSR position (aesc.h:392-412, 508-521)
  float smx = 0.0f, smn = FLT_MAX;
  for (uint32_t p = r0 + threadIdx.x; p < r1; p += kPB) {
    const float x = pv[A.adj[p] - off];
    smx = fmaxf(smx, x);
    smn = fminf(smn, x);
  }
  smx = block_reduce<kPW>(smx, sh.rf, OpMaxF{});
  smn = block_reduce<kPW>(smn, sh.rf, OpMinF{});
  const float dse = static_cast<float>(dse_d);
  auto emit_nr = [&](uint32_t p, int rest, float tmx, float tmn) {
    const int len = rest;
    const float chi = gmax + (smx + tmx) / 2.0 + (len - 1) * emax - smn - tmn -
                      2 * (len - 1) * gmin;
    const double v = fmax(ceil(8 * (chi * chi) * logmd / (dse * dse)), 1.0);
    uint64_t nr = 0;
    if (!(v < kMaxWalksFast)) {
      atomicOr(A.status, 1u);
    } else {
      nr = static_cast<uint64_t>(v);
      nr += 8 - nr % 8;
    }
    A.nr_pos[p] = nr;
  };
  if (rs.y > 64) {
    // a hub source: one 8-lane group per position, 32 positions at a time
    // (a wave per position made its positions the workgroup's longest
    // chain); max and min are exact in any order
    const uint32_t cg = threadIdx.x >> 3, cb = threadIdx.x & 7u;
    for (uint32_t p = r0 + cg; p < r1; p += kPB / 8) {
      const int rest = A.taus[p] - ell;
      if (rest <= 0) {
        if (cb == 0) A.nr_pos[p] = 0;
        continue;
      }
      const uint2 rv = A.rowd[A.adj[p]];
      float tmx = 0.0f, tmn = FLT_MAX;
      for (uint32_t k = cb; k < rv.y; k += 8) {
        const float x = pv[A.adj[rv.x + k] - off];
        tmx = fmaxf(tmx, x);
        tmn = fminf(tmn, x);
      }
#pragma unroll
      for (int o = 4; o > 0; o >>= 1) {
        tmx = fmaxf(tmx, __shfl_xor(tmx, o));
        tmn = fminf(tmn, __shfl_xor(tmn, o));
      }
      if (cb == 0) emit_nr(p, rest, tmx, tmn);
    }
  } else {
    // one wave per position: the neighbour's row in 64-wide steps
    for (uint32_t p = r0 + wave_id(); p < r1; p += kPW) {
      const int rest = A.taus[p] - ell;
      if (rest <= 0) {
        if (lane_id() == 0) A.nr_pos[p] = 0;
        continue;
      }
      const uint2 rv = A.rowd[A.adj[p]];
      float tmx = 0.0f, tmn = FLT_MAX;
      for (uint32_t k = lane_id(); k < rv.y; k += 64) {
        const float x = pv[A.adj[rv.x + k] - off];
        tmx = fmaxf(tmx, x);
        tmn = fminf(tmn, x);
      }
      tmx = wave_max(tmx);
      tmn = wave_min(tmn);
      if (lane_id() == 0) emit_nr(p, rest, tmx, tmn);
    }
  }
  __syncthreads();

  // ---- 2-way walks.  Positions with at most small_bq bouquets: one 8-lane
  // group each (32 groups side by side); larger ones: the whole workgroup.
  const uint32_t g = threadIdx.x >> 3, b = threadIdx.x & 7u;
  const uint32_t giw = g & 7u, gbase = (threadIdx.x & 63u) & ~7u;
  const uint64_t small_bq = A.small_bq;
  for (uint32_t p = r0 + g; p < r1; p += kPB / 8) {
    const int rest = A.taus[p] - ell;
    const uint64_t nr = rest > 0 ? A.nr_pos[p] : 0;
    if (nr == 0 || nr / 8 > small_bq) continue;
    const uint32_t v = A.adj[p];
    const uint32_t step = static_cast<uint32_t>(0xFFFFFFFFull / nr);
    // bouquet sums grouped by bouquet index mod 8 and the eight partial
    // sums added in order: the order of the eight-group path below, so a
    // position's sum does not depend on small_bq (which a batch autotunes)
    const uint32_t nbq = static_cast<uint32_t>(nr / 8);
    double part[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t j0 = 0; j0 < nbq; j0 += 8) {
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k)
        if (j0 + k < nbq)
          part[k] += group_sum8(walk2_bouquet(A, pv, off, src, v, j0 + k, b, step, rest, giw),
                                gbase);
    }
    double acc = 0.0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) acc += part[k];
    if (b == 0) {
      A.preds_d[p] += acc / (dsd * static_cast<double>(nr));
      if (A.tr_x) A.tr_x[p] = static_cast<float>(acc);
    }
  }
  __syncthreads();
  // Larger positions: four at a time, eight groups each.  Position slot i
  // (groups 8i .. 8i+7) takes bouquets j = g % 8 + 8t of its position; its
  // sum is the eight group sums in group order.  The list of larger
  // positions is built in row order, 256 positions at a time.
  for (uint32_t base = r0; base < r1; base += kPB) {
    const uint32_t p0 = base + threadIdx.x;
    bool big = false;
    if (p0 < r1) {
      const int rest = A.taus[p0] - ell;
      const uint64_t nr = rest > 0 ? A.nr_pos[p0] : 0;
      big = nr / 8 > small_bq;
    }
    uint32_t nbig = 0;
    const uint32_t at = block_exclusive_scan<kPW>(big ? 1u : 0u, sh.ru, &nbig);
    if (big) sh.heavy[at] = p0;
    __syncthreads();
    for (uint32_t q = 0; q < nbig; q += 4) {
      const uint32_t slot = g >> 3, gi = g & 7u;
      double acc = 0.0;
      if (q + slot < nbig) {
        const uint32_t p = sh.heavy[q + slot];
        const int rest = A.taus[p] - ell;
        const uint64_t nr = A.nr_pos[p];
        const uint32_t v = A.adj[p];
        const uint32_t step = static_cast<uint32_t>(0xFFFFFFFFull / nr);
        const uint64_t nbq = nr / 8;
        for (uint64_t jb = gi; jb < nbq; jb += 8)
          acc += group_sum8(walk2_bouquet(A, pv, off, src, v, static_cast<uint32_t>(jb), b, step,
                                          rest, giw),
                            gbase);
      }
      if (b == 0) sh.wacc[g] = acc;
      __syncthreads();
      if (threadIdx.x < 4 && q + threadIdx.x < nbig) {
        const uint32_t pp = sh.heavy[q + threadIdx.x];
        double sum = 0.0;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) sum += sh.wacc[8 * threadIdx.x + k];
        const uint64_t nrr = A.nr_pos[pp];
        A.preds_d[pp] += sum / (dsd * static_cast<double>(nrr));
        if (A.tr_x) A.tr_x[pp] = static_cast<float>(sum);
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    if (A.tr_ell) A.tr_ell[src] = ell;
    if (A.tr_emax) A.tr_emax[src] = emax;
    if (A.tr_gmax) A.tr_gmax[src] = gmax;
    if (A.tr_gmin) A.tr_gmin[src] = gmin;
    if (A.tr_qsize) A.tr_qsize[src] = Q;
  }
  __syncthreads();
}

// Vectors of blocks with n <= stage_max (the floats of dynamic LDS the
// launch provides, at most kPostLds) are staged in LDS; larger ones are
// read in place.
__global__ void __launch_bounds__(kPB) blk_post_kernel(BlockArgs A, uint32_t b0, uint32_t b1,
                                                       uint32_t stage_max) {
  extern __shared__ float lds_pv[];
  __shared__ float cand[kPB * kTopK];
  __shared__ PostShared sh;
  const uint32_t total = (b1 - b0) * kBlkCols;
  for (;;) {
    if (threadIdx.x == 0) sh.k = atomicAdd(A.next_src, 1u);
    __syncthreads();
    const uint32_t k = sh.k;
    __syncthreads();
    if (k >= total) break;
    const uint32_t b = b0 + k / kBlkCols, j = k % kBlkCols;
    const BlockInfo B = A.blocks[b];
    if (j >= B.cnt) continue;  // uniform per workgroup
    const uint32_t L = A.last[b];
    const float* O = ((L & 1u) ? A.xa : A.xb) + B.xoff + static_cast<uint64_t>(j) * B.n;
    const float* pv = O;
    if (B.n <= stage_max) {
      for (uint32_t u0 = threadIdx.x; u0 < B.n; u0 += kPB * kPassU) {
        float xs[kPassU];
#pragma unroll
        for (uint32_t c = 0; c < kPassU; ++c) {
          const uint32_t u = u0 + c * kPB;
          xs[c] = u < B.n ? O[u] : 0.0f;
        }
#pragma unroll
        for (uint32_t c = 0; c < kPassU; ++c)
          if (u0 + c * kPB < B.n) lds_pv[u0 + c * kPB] = xs[c];
      }
      __syncthreads();
      pv = lds_pv;
    }
    post_source(A, B, b * kBlkCols + j, B.src0 + j, pv, cand, sh);
  }
}

}  // namespace

int block_post_per_cu(uint32_t stage) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void*>(&blk_post_kernel), kPB,
          static_cast<size_t>(stage) * sizeof(float)) != hipSuccess)
    per_cu = 1;
  return std::max(1, per_cu);
}

int LaunchTimer::begin(hipStream_t s, int k) {
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess) return -1;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return -1;
  }
  ev.push_back(a);
  ev.push_back(b);
  kind.push_back(k);
  (void)hipEventRecord(a, s);
  return static_cast<int>(kind.size()) - 1;
}

void LaunchTimer::end(hipStream_t s, int pair) {
  if (pair >= 0) (void)hipEventRecord(ev[2 * pair + 1], s);
}

LaunchTimer::~LaunchTimer() {
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
}

int launch_block_subset(tgtp_ctx* ctx, hipStream_t s, const BlockArgs& A, const BlockSubset& S,
                        const uint32_t* round_tiles, const uint32_t* round_hubs,
                        uint32_t post_slots, LaunchTimer* timer) {
  auto tb = [&](int k) { return timer ? timer->begin(s, k) : -1; };
  auto te = [&](int pair) {
    if (timer) timer->end(s, pair);
  };
  const uint32_t nb = S.b1 - S.b0;
  const uint32_t col_blocks = (nb * kBlkCols + 3) / 4;
  const uint64_t zb = std::min<uint64_t>(16384, (S.floats + 255) / 256 + 1);
  blk_zero_kernel<<<static_cast<uint32_t>(zb), 256, 0, s>>>(A.xb, S.floats);
  blk_init_kernel<<<col_blocks, 256, 0, s>>>(A, S.b0, S.b1);
  blk_first_kernel<<<col_blocks, 256, 0, s>>>(A, S.b0, S.b1);
  blk_stop_kernel<<<col_blocks, 256, 0, s>>>(A, S.b0, S.b1, A.xb);
  // tile units: tiles (mode 0, grid rounded up to a multiple of 8) or
  // groups of 8 tiles (mode 1)
  auto grid_for = [&](uint32_t units) { return A.tile_mode == 0 ? (units + 7) / 8 * 8 : 8 * units; };
  const uint32_t tile_grid = grid_for(S.tiles);
  for (uint32_t r = 2; r <= S.rounds; ++r) {
    const float* cur = (r & 1u) ? A.xa : A.xb;
    float* nxt = (r & 1u) ? A.xb : A.xa;
    const uint32_t tr = round_tiles[r];
    if (tr == 0) continue;
    const uint32_t hr = round_hubs[r], hub_wgs = (hr + 7u) & ~7u;
    const int tr_pair = tb(LaunchTimer::kRound);
    if (A.round_form == 1) {
      blk_round4_kernel<8, 5><<<grid_for(tr), 256, 0, s>>>(A, S.b0, S.b1, tr, 0u, cur, nxt);
      if (hr) blk_hub_kernel<<<hr, 256, 0, s>>>(A, S.b0, S.b1, hr, cur, nxt);
    } else if (A.round_form == 2)
      blk_round4_kernel<16, 4><<<hub_wgs + grid_for(tr), 256, 0, s>>>(A, S.b0, S.b1, tr, hr, cur,
                                                                       nxt);
    else {
      blk_round_kernel<<<grid_for(tr), 256, 0, s>>>(A, S.b0, S.b1, tr, cur, nxt);
      if (hr) blk_hub_kernel<<<hr, 256, 0, s>>>(A, S.b0, S.b1, hr, cur, nxt);
    }
    te(tr_pair);
    const int st_pair = tb(LaunchTimer::kStop);
    blk_stop_kernel<<<col_blocks, 256, 0, s>>>(A, S.b0, S.b1, nxt);
    te(st_pair);
  }
  blk_transpose_kernel<<<tile_grid, 256, 0, s>>>(A, S.b0, S.b1, S.tiles);
  blk_zero_kernel<<<1, 256, 0, s>>>(reinterpret_cast<float*>(A.next_src), 1);
  // the LDS buffer holds the subset's largest staged vector (smaller
  // buffers, more workgroups per CU; none when every vector is read in place)
  const int po_pair = tb(LaunchTimer::kPost);
  blk_post_kernel<<<std::max(1u, std::min(post_slots, nb * kBlkCols)), kPB,
                    static_cast<size_t>(S.post_stage) * sizeof(float), s>>>(A, S.b0, S.b1,
                                                                           S.post_stage);
  te(po_pair);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}
