// saba_part.hip -- SABA 1-way walks on large graphs, hop-synchronous over
// vertex partitions.
//
// Semantics are those of saba_walk.hip (rw_experiment_saba.cpp:65-145): walk
// i = 8*j + b of source s starts at s with secret (i+1)*floor((2^32-1)/W);
// each hop picks k = hi32((sec ^ fmix32(cur)) * deg) and multiplies sec by
// deg, until some lane of the bouquet holds sec == 0 at the start of a hop,
// after which all eight lanes draw k = splitmix32(state) % deg; the visit of
// every vertex reached is counted.  A walker at a vertex without neighbours
// stops there.
//
// Why a second engine: the direct kernel follows each walker through all its
// hops, so every hop is one random cache line fetched from the whole
// adjacency (~84 MB at n = 2^20, m = 2*10^7) -- the rate of random lines the
// memory system serves bounds it.  Here a chunk of walkers advances one hop
// per launch ("round").  Between rounds each walker is a 12-byte record
// {vertex, secret, walker id} in the bucket of the partition holding its
// vertex (8 bytes {vertex, secret} into the round that takes the last hop,
// which marks no bouquet, and 4 bytes, the vertex, into the last round, which
// only counts arrivals); partitions are contiguous vertex ranges whose rows
// and adjacency (~0.7 MB) fit an XCD's L2.  Each partition is dealt to one XCD (heaviest
// first by degree sum, the rate at which mixed walks visit it) and workgroup
// b serves, in increasing order, the partitions dealt to XCD b % 8 --
// workgroups are dealt round-robin to the eight XCDs, so each XCD's L2 holds
// the partitions its workgroups are on -- and every gather of a round is an
// L2 hit, while the records stream through HBM once per hop in coalesced
// runs.
//
// Records of one round are written in blocks of kBR records, each workgroup
// taking blocks in order from a region of the pool sized to its input (no
// atomics); a workgroup sorts each tile of up to kBR outgoing records by
// destination partition in LDS and appends each run to its current block of
// that partition.  A small kernel then lists every partition's blocks for
// the next round.  The pools of a pass take up to 96 GiB: fewer, larger
// source chunks, and each round's fixed share (a partial block per
// partition and workgroup, the launch's tail) paid fewer times.
//
// Visits are counted when a record is read (the walker has arrived): in LDS
// counters of the partition, flushed per (workgroup, partition) by plain
// read-add-write into the workgroup's own copy of the counter vector
// (workgroups of one round never share a (copy, vertex) pair); the copies are
// summed into the caller's counters at the end.
//
// Bouquet fallback: lanes of one bouquet are routed to different partitions,
// so a round cannot test the bouquet-wide zero.  Every lane walks with its
// own secret; a lane whose secret becomes zero while more hops follow (or
// starts at zero) marks its bouquet in a bitmap.  After the last round a
// fix-up kernel replays each marked bouquet in full: it finds the first hop
// at which the bouquet would have switched, then for every lane subtracts the
// visits of the independent path from that hop on and adds those of the
// SplitMix32 path.  Counts are integers, so the result is exact.  On the
// n = 2^20 walk workload about 1 bouquet in 2,000 is marked.
#include "common.hpp"

#include <algorithm>
#include <cstdlib>
#include <vector>

namespace {

int env_int(const char* name, int dflt) {
  const char* s = std::getenv(name);
  return s && *s ? std::atoi(s) : dflt;
}

#ifndef TGTP_PART_PER
#define TGTP_PART_PER 6
#endif
#ifndef TGTP_PART_CNT16
#define TGTP_PART_CNT16 1
#endif
constexpr int kPBlock = 1024;                // threads per workgroup
constexpr uint32_t kPer = TGTP_PART_PER;     // records per thread per tile
constexpr uint32_t kBR = kPBlock * kPer;     // records per block = per tile
constexpr uint32_t kMaxParts = 256;          // partitions (= sort bins) at most
constexpr uint32_t kTab = 2048;              // coarse vertex -> partition table
constexpr uint32_t kList = 512;              // block ids of a partition staged in LDS at a time
// Arrival counters of the partition a workgroup serves, in LDS: 16 bits per
// vertex (two to a word, added into the workgroup's copy every kFlushTiles
// tiles, which is at most 65,535 arrivals, so no half carries into the other)
// or 32.  At 2 bytes a vertex, partitions of 10,240 vertices fit beside the
// tile buffers in 160 KiB of LDS (8,192 at 4 bytes): fewer partitions, fewer
// partial blocks per round.
constexpr bool kCnt16 = TGTP_PART_CNT16 != 0;
constexpr uint32_t kFlushTiles = 65535u / kBR;
static_assert(kFlushTiles >= 1, "a tile's arrivals fit 16 bits");
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kMaxPoolBlocks = 0xFFFFFFFFu / kBR;  // record index fits 32 bits

enum Mode { kGen = 0, kMid = 1, kLast = 2 };

// A round's output pool: block b holds bfill[b] records of partition
// bpart[b], the bslot[b]-th block that workgroup bwg[b] wrote for it.  Each
// workgroup w writes blocks [reg_start[w], reg_start[w] + reg_used[w]) of a
// region sized to its input (a workgroup writes at most one record per record
// it reads: its full blocks number at most its input blocks, plus one partial
// block per partition), so no block is taken with a global atomic.
struct PartRound {
  const uint2* rowd;
  const uint32_t* idx;
  const uint32_t* adj;
  const uint32_t* plo;   // [P+1] partition vertex bounds
  const uint16_t* ptab;  // [kTab+2]
  const uint32_t* tadj;  // adjacency tagged with partitions (vertex | p << 24), or null
  const uint8_t* pxcd;   // [kMaxParts] workgroups b with b % 8 == pxcd[p] serve partition p
  uint64_t n;
  uint32_t P;
  uint32_t vbig;         // vertices of the largest partition
  uint32_t tshift;
  uint32_t depth;
  uint32_t round;        // hop performed by this launch (0-based); kLast: depth
  uint32_t W;
  uint32_t step;
  uint32_t src_first;    // source of chunk walker id w: src_first + (w / W) * src_stride
  uint32_t src_stride;
  uint32_t n_walkers;    // walkers of the chunk (kGen)
  uint32_t pool_cap;     // blocks in the output pool
  // input pool (kMid, kLast)
  const uint32_t* rec_in;     // [blocks][kBR][3] = {vertex, secret, walker id},
                              // [..][2] {vertex, secret} or [..][1] vertex
  const uint32_t* order_in;   // blocks grouped by partition
  const uint32_t* pstart_in;  // [P] first entry of each partition in order_in
  const uint32_t* bcount_in;  // [P] blocks per partition
  const uint32_t* bfill_in;   // [blocks] records per block
  // output pool (kGen, kMid)
  uint32_t* rec_out;
  uint32_t* bfill_out;
  uint32_t* bpart_out;
  uint32_t* bslot_out;
  uint32_t* cwd;              // [P][grid] blocks written per (partition, workgroup)
  uint32_t* reg_start;        // [grid]
  uint32_t* reg_used;         // [grid]
  uint32_t* overflow;         // set when a region runs out (never with sized regions)
  uint32_t* copies;           // [grid / 8][n]
  uint32_t* counters;         // the caller's counters (arrivals of a two-hop generation round)
  uint32_t gen_hops;          // hops the generation round takes (at most depth)
  uint32_t gen_slots;         // two-hop generation, W >= tile: first-hop arrivals counted
                              // per neighbour slot of the tile's (at most two) sources
  uint32_t* taint;            // bouquet bitmap of the chunk
};

struct PartLds {
  uint32_t hist[kMaxParts];
  uint32_t off[kMaxParts];
  uint4 run[kMaxParts];
  uint32_t curblk[kMaxParts];
  uint32_t curfill[kMaxParts];
  uint32_t lcount[kMaxParts]; // blocks this workgroup wrote per partition
  uint32_t wsum[16];
  uint32_t total;
  uint32_t rbase;             // first block of this workgroup's region
  uint32_t rsize;             // blocks in it
  uint32_t used;              // blocks taken from it
  uint32_t hn;                // generation round: entries in the arrival table
  uint32_t plo[kMaxParts + 1];
  uint16_t ptab[kTab + 2];
  uint8_t pxcd[kMaxParts];
  uint32_t lblk[kList];       // this workgroup's input blocks of the current partition
  uint32_t lfill[kList];
  uint32_t tcur[kBR];
  uint32_t tsec[kBR];
  uint32_t twid[kBR];
  uint8_t tdest[kBR];
};

// Workgroup barrier ordering LDS only: global loads issued before it stay in
// flight (a prefetched tile), which a full __syncthreads() would drain.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ uint32_t part_of(const PartLds& L, uint32_t v, uint32_t tshift) {
  const uint32_t t = v >> tshift;
  uint32_t a = L.ptab[t], b = L.ptab[t + 1];
  while (a < b) {  // largest p in [a, b] with plo[p] <= v
    const uint32_t mid = (a + b + 1u) >> 1;
    if (L.plo[mid] <= v) a = mid;
    else b = mid - 1u;
  }
  return a;
}

__device__ __forceinline__ void mark_bouquet(uint32_t* taint, uint32_t wid) {
  const uint32_t bq = wid >> 3;
  atomicOr(&taint[bq >> 5], 1u << (bq & 31u));
}

// Sum of v over threads [0, 256) (others pass 0); every thread of the block
// calls it and gets the sum.
__device__ __forceinline__ uint32_t sum256(PartLds& L, uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (threadIdx.x < kMaxParts && (threadIdx.x & 63u) == 0u) L.wsum[threadIdx.x >> 6] = v;
  lds_barrier();
  const uint32_t r = L.wsum[0] + L.wsum[1] + L.wsum[2] + L.wsum[3];
  lds_barrier();
  return r;
}

// Exclusive scan of hist[0, 256) into off[], sum into total.  Every thread of
// the block calls it; it ends with a barrier.
__device__ __forceinline__ void scan_bins(PartLds& L) {
  const uint32_t tid = threadIdx.x;
  uint32_t v = 0, incl = 0;
  if (tid < kMaxParts) {
    v = L.hist[tid];
    incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o);
      if ((tid & 63u) >= static_cast<uint32_t>(o)) incl += y;
    }
    if ((tid & 63u) == 63u) L.wsum[tid >> 6] = incl;
  }
  lds_barrier();
  if (tid < kMaxParts) {
    uint32_t add = 0;
    for (uint32_t w = 0; w < (tid >> 6); ++w) add += L.wsum[w];
    L.off[tid] = incl - v + add;
    if (tid == kMaxParts - 1) L.total = incl + add;
  }
  lds_barrier();
}

// Counts one arrival at v in the LDS table (open addressing, H a power of two
// or 0), or directly in the counters when the table has no slot for it.
__device__ __forceinline__ void count_arrival(PartLds& L, uint32_t* hkey, uint32_t* hval,
                                              uint32_t* hlist, uint32_t H, uint32_t v,
                                              uint32_t* counters) {
  if (H) {
    uint32_t h = fmix32(v) & (H - 1u);
    for (int probe = 0; probe < 8; ++probe) {
      const uint32_t k = atomicCAS(&hkey[h], kNone, v);
      if (k == kNone || k == v) {
        atomicAdd(&hval[h], 1u);
        if (k == kNone) hlist[atomicAdd(&L.hn, 1u)] = h;
        return;
      }
      h = (h + 1u) & (H - 1u);
    }
  }
  atomicAdd(&counters[v], 1u);
}

// One tile: each thread holds up to kPer walkers at vertex cur[j] (kNone: none)
// that take hop `round`.  Sorts the successors by partition and appends them
// to the workgroup's blocks.  kLdsRows: the walkers are in the partition whose
// row offsets ridx[0, V] (vertices lo + i) are staged in LDS, so the hop's only
// gather is the neighbour itself; otherwise the row descriptors are gathered.
// Writes the sorted tile held in LDS to its blocks.
//
// Every thread issues the same kPer stores whatever the tile's size: slots past
// it write to the pool's spare block (block pool_cap, never read as records),
// so no branch or loop of data-dependent length stands between the
// neighbour gathers and their use, and the compiler's wait before that use
// counts exactly the loads and stores issued after the gathers instead of
// waiting for every outstanding access.
//
// kOutW: 32-bit words per record written.  3: {vertex, secret, walker id}.
// 2: {vertex, secret}, for the round that takes the last hop (a bouquet is
// marked only while more hops follow, so it needs no walker id).  1: the
// vertex alone, for the last round, which only counts arrivals.  Record g
// is at word kOutW * g.
template <int kOutW>
__device__ __forceinline__ void copy_out(PartLds& L, const PartRound& A) {
  const uint32_t total = L.total;
  const uint64_t spare = static_cast<uint64_t>(A.pool_cap) * kBR + threadIdx.x;
#pragma unroll
  for (uint32_t j = 0; j < kPer; ++j) {
    const uint32_t q = j * kPBlock + threadIdx.x;
    const uint4 rn = L.run[L.tdest[q]];
    const uint32_t g = q + (q < rn.x ? rn.y : rn.z);
    const uint64_t e = q < total ? static_cast<uint64_t>(g) : spare;
    uint32_t* r = A.rec_out + e * static_cast<uint32_t>(kOutW);
    __builtin_nontemporal_store(L.tcur[q], r);
    if constexpr (kOutW >= 2) __builtin_nontemporal_store(L.tsec[q], r + 1);
    if constexpr (kOutW >= 3) __builtin_nontemporal_store(L.twid[q], r + 2);
  }
}

// `pending`: LDS holds the previous tile, sorted; its records are written
// after this tile's neighbour gathers are issued.  Vector-memory completions
// are counted in issue order, so stores issued before the gathers would hold
// up the wait for them.  The caller writes the last tile (copy_out).  (The
// next tile's record loads are issued at the top of the tile, before the
// gathers; issued after them -- the next tile's or the one after's -- the
// pass was 8-10% slower.)
//
// `arrive` (a middle round) counts the tile's arrivals in the partition's LDS
// counters once the gathers are issued: LDS operations complete in issue
// order, so counting first would hold up the row offsets the gathers need.
template <bool kLdsRows, int kOutW, typename Arrive>
__device__ __forceinline__ void step_and_scatter(PartLds& L, const PartRound& A, uint32_t hop,
                                                 const uint32_t* ridx, uint32_t lo,
                                                 uint32_t (&cur)[kPer], uint32_t (&sec)[kPer],
                                                 uint32_t (&wid)[kPer], bool& pending,
                                                 Arrive&& arrive) {
  const uint32_t tid = threadIdx.x;
  bool valid[kPer];  // kNone marks an empty slot
#pragma unroll
  for (uint32_t j = 0; j < kPer; ++j) valid[j] = cur[j] != kNone;
  // (branch-free: empty slots read row 0 / entry 0, and their results are
  // dropped; a bouquet to mark is marked once the gathers have been used)
  uint2 rd[kPer];
#pragma unroll
  for (uint32_t j = 0; j < kPer; ++j) {
    if constexpr (kLdsRows) {
      const uint32_t o = valid[j] ? cur[j] - lo : 0u;
      const uint32_t a = ridx[o];
      rd[j] = make_uint2(a, ridx[o + 1u] - a);
    } else {
      rd[j] = A.rowd[valid[j] ? cur[j] : 0u];
    }
  }
  uint32_t nxt[kPer], dest[kPer], rank[kPer];
  bool mark[kPer];
  const uint32_t* gadj = A.tadj ? A.tadj : A.adj;
#pragma unroll
  for (uint32_t j = 0; j < kPer; ++j) {
    valid[j] = valid[j] && rd[j].y != 0u;
    const uint32_t k = __umulhi(sec[j] ^ fmix32(cur[j]), rd[j].y);
    const uint32_t ns = sec[j] * rd[j].y;
    mark[j] = valid[j] && ns == 0u && sec[j] != 0u && hop + 1u < A.depth;
    sec[j] = valid[j] ? ns : sec[j];
    nxt[j] = gadj[valid[j] ? rd[j].x + k : 0u];
  }
  // keeps the gathers ahead of the stores in issue order
  __asm__ volatile("" ::: "memory");
  arrive();
  copy_out<kOutW>(L, A);  // the previous tile (none yet: L.total is 0)
  if (A.tadj) {  // uniform
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
      dest[j] = nxt[j] >> 24;
      nxt[j] &= 0xFFFFFFu;
    }
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) dest[j] = valid[j] ? part_of(L, nxt[j], A.tshift) : 0u;
  }
#pragma unroll
  for (uint32_t j = 0; j < kPer; ++j)
    if (mark[j]) mark_bouquet(A.taint, wid[j]);  // rare
#pragma unroll
  for (uint32_t j = 0; j < kPer; ++j) rank[j] = valid[j] ? atomicAdd(&L.hist[dest[j]], 1u) : 0u;
  lds_barrier();
  scan_bins(L);
  if (tid < A.P) {
    const uint32_t c = L.hist[tid];
    uint32_t cb = L.curblk[tid], cf = L.curfill[tid];
    const uint32_t l0 = min(kBR - cf, c);
    const uint32_t o = L.off[tid];
    const uint32_t seg0 = cb == kNone ? 0u : cb * kBR + cf;
    uint32_t seg1 = A.pool_cap * kBR;
    if (c > l0) {
      if (cb != kNone) A.bfill_out[cb] = kBR;  // full
      const uint32_t i = atomicAdd(&L.used, 1u);
      const uint32_t nb = L.rbase + i;
      if (i < L.rsize && nb < A.pool_cap) {
        A.bpart_out[nb] = tid;
        A.bslot_out[nb] = L.lcount[tid]++;
        seg1 = nb * kBR;
        cb = nb;
        cf = c - l0;
      } else {
        atomicOr(A.overflow, 1u);
        cb = kNone;
        cf = kBR;
      }
    } else {
      cf += l0;
    }
    L.run[tid] = make_uint4(o + l0, seg0 - o, seg1 - o - l0, 0u);
    L.curblk[tid] = cb;
    L.curfill[tid] = cf;
  }
  if (tid < kMaxParts) L.hist[tid] = 0u;  // read above by this thread only
#pragma unroll
  for (uint32_t j = 0; j < kPer; ++j) {
    if (valid[j]) {
      const uint32_t pos = L.off[dest[j]] + rank[j];
      L.tcur[pos] = nxt[j];
      if constexpr (kOutW >= 2) L.tsec[pos] = sec[j];
      if constexpr (kOutW >= 3) L.twid[pos] = wid[j];
      L.tdest[pos] = static_cast<uint8_t>(dest[j]);
    }
  }
  lds_barrier();
  pending = true;
}

// Loads input record i of block blk (kNone: none) as vertex / secret / id;
// slots past the block's fill get vertex kNone (branch-free: they read the
// pool's spare block, and the values are dropped).  kInW: words per record
// (as copy_out's kOutW); fields the records do not hold read as 0.
template <int kInW>
__device__ __forceinline__ void load_tile(const PartRound& A, uint32_t blk, uint32_t fill,
                                          uint32_t (&cur)[kPer], uint32_t (&sec)[kPer],
                                          uint32_t (&wid)[kPer]) {
#pragma unroll
  for (uint32_t j = 0; j < kPer; ++j) {
    const uint32_t i = j * kPBlock + threadIdx.x;
    const bool ok = blk != kNone && i < fill;
    const uint64_t g = ok ? static_cast<uint64_t>(blk) * kBR + i
                          : static_cast<uint64_t>(A.pool_cap) * kBR + threadIdx.x;
    const uint32_t* r = A.rec_in + g * static_cast<uint32_t>(kInW);
    const uint32_t c = __builtin_nontemporal_load(r);
    cur[j] = ok ? c : kNone;
    if constexpr (kInW >= 2) {
      const uint32_t x = __builtin_nontemporal_load(r + 1);
      sec[j] = ok ? x : 0u;
    } else {
      sec[j] = 0u;
    }
    if constexpr (kInW >= 3) {
      const uint32_t y = __builtin_nontemporal_load(r + 2);
      wid[j] = ok ? y : 0u;
    } else {
      wid[j] = 0u;
    }
  }
}

// kInW / kOutW: words per record read (middle and last rounds) and written
// (generation and middle rounds), as copy_out's kOutW.
template <int kMode, int kInW, int kOutW>
__global__ void __launch_bounds__(kPBlock)
part_round_kernel(PartRound A) {
  static_assert((kMode == kLast) == (kOutW == 0), "the last round writes no records");
  static_assert(kMode != kGen || kInW == 0, "the generation round reads no records");
  extern __shared__ __align__(16) uint8_t lds_raw[];
  PartLds& L = *reinterpret_cast<PartLds*>(lds_raw);
  // the current partition's arrival counters and row offsets
  uint32_t* cnt = reinterpret_cast<uint32_t*>(lds_raw + sizeof(PartLds));
  const uint32_t cnt_words = kCnt16 ? (A.vbig + 1u) / 2u : A.vbig;
  uint32_t* ridx = cnt + cnt_words;
  const uint32_t area = cnt_words + A.vbig + 1u;  // words after PartLds
  const uint32_t tid = threadIdx.x;
  const uint32_t w = blockIdx.x;
  const uint32_t xc = w & 7u;
  const uint32_t l = w >> 3;
  const uint32_t NL = gridDim.x >> 3;
  for (uint32_t i = tid; i <= A.P; i += kPBlock) L.plo[i] = A.plo[i];
  for (uint32_t i = tid; i < kTab + 2; i += kPBlock) L.ptab[i] = A.ptab[i];
  if (tid < kMaxParts) L.pxcd[tid] = A.pxcd[tid];
  if (tid < kMaxParts) {
    L.hist[tid] = 0u;
    L.curblk[tid] = kNone;
    L.curfill[tid] = kBR;
    L.lcount[tid] = 0u;
  }
  if (tid == 0) {
    L.used = 0u;
    L.total = 0u;  // no tile to write out yet
  }
  // this workgroup's output region: the input blocks of workgroups before it
  // plus P each, then its own input blocks plus P
  if constexpr (kMode == kGen) {
    const uint32_t n_tiles = (A.n_walkers + kBR - 1u) / kBR;
    const uint32_t q = n_tiles / gridDim.x, rem = n_tiles % gridDim.x;
    if (tid == 0) {
      L.rbase = q * w + min(w, rem) + w * A.P;
      L.rsize = q + (w < rem ? 1u : 0u) + A.P;
    }
    lds_barrier();
  } else if constexpr (kMode == kMid) {
    uint32_t below = 0, mine = 0;
    if (tid < A.P) {
      const uint32_t nbp = A.bcount_in[tid];
      const uint32_t xp = L.pxcd[tid];
      const uint32_t Lp = w > xp ? (w - xp + 7u) >> 3 : 0u;  // l' with 8 l' + xp < w
      below = (nbp / NL) * Lp + min(nbp % NL, Lp);
      if (xp == xc) mine = nbp / NL + (l < nbp % NL ? 1u : 0u);
    }
    lds_barrier();
    below = sum256(L, below);
    mine = sum256(L, mine);
    if (tid == 0) {
      L.rbase = below + w * A.P;
      L.rsize = mine + A.P;
    }
    lds_barrier();
  } else {
    lds_barrier();
  }

  bool pending = false;  // LDS holds a sorted tile not yet written
  if constexpr (kMode == kGen) {
    // arrival table of a two-hop generation round, in the space the other
    // rounds use for counters and row offsets: H slots of key, count and
    // the list of slots in use
    uint32_t H = 0;
    if (A.gen_hops > 1u) {
      const uint32_t words = area;
      while ((H ? H << 1 : 1u) * 3u <= words && H < (1u << 16)) H = H ? H << 1 : 1u;
      if (H < 64u) H = 0;  // too small to help: every arrival goes to the counters
    }
    uint32_t* hkey = cnt;
    uint32_t* hval = cnt + H;
    uint32_t* hlist = cnt + 2u * H;
    if (A.gen_slots) {
      for (uint32_t i = tid; i < area; i += kPBlock) cnt[i] = 0u;
    } else {
      for (uint32_t i = tid; i < H; i += kPBlock) {
        hkey[i] = kNone;
        hval[i] = 0u;
      }
    }
    if (tid == 0) L.hn = 0u;
    lds_barrier();
    const uint32_t n_tiles = (A.n_walkers + kBR - 1u) / kBR;
    for (uint32_t t = w; t < n_tiles; t += gridDim.x) {
      uint32_t cur[kPer], sec[kPer], wid[kPer];
      bool valid[kPer];
#pragma unroll
      for (uint32_t j = 0; j < kPer; ++j) {
        wid[j] = t * kBR + j * kPBlock + tid;
        valid[j] = wid[j] < A.n_walkers;
        const uint32_t sl = valid[j] ? wid[j] / A.W : 0u;
        const uint32_t i = valid[j] ? wid[j] - sl * A.W : 0u;
        cur[j] = A.src_first + sl * A.src_stride;
        sec[j] = (i + 1u) * A.step;
        if (valid[j] && sec[j] == 0u) mark_bouquet(A.taint, wid[j]);
      }
      if (A.gen_slots) {
        // the first hop here, arrivals counted per neighbour slot of the
        // tile's sources (uniform: s0 holds wids from the tile's first on, s1
        // the rest); the second hop reads the rows of those neighbours from
        // cache in step_and_scatter
        const uint32_t w0 = t * kBR, w1 = min(w0 + kBR, A.n_walkers) - 1u;
        const uint32_t sl0 = w0 / A.W, sl1 = w1 / A.W;
        const uint2 r0 = A.rowd[A.src_first + sl0 * A.src_stride];
        const uint2 r1 = A.rowd[A.src_first + sl1 * A.src_stride];
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
          const bool second = valid[j] && wid[j] / A.W != sl0;
          const uint2 rd = second ? r1 : r0;
          valid[j] = valid[j] && rd.y != 0u;
          if (valid[j]) {
            const uint32_t k = __umulhi(sec[j] ^ fmix32(cur[j]), rd.y);
            const uint32_t ns = sec[j] * rd.y;
            if (ns == 0u && sec[j] != 0u) mark_bouquet(A.taint, wid[j]);  // depth >= 2
            sec[j] = ns;
            cur[j] = A.adj[rd.x + k];
            atomicAdd(&cnt[(second ? r0.y : 0u) + k], 1u);
          }
        }
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) cur[j] = valid[j] ? cur[j] : kNone;
        step_and_scatter<false, kOutW>(L, A, 1u, nullptr, 0u, cur, sec, wid, pending, [] {});
        // the barrier at the end of step_and_scatter follows every count
        const uint32_t ns0 = r0.y, nsl = r0.y + (sl1 != sl0 ? r1.y : 0u);
        for (uint32_t i = tid; i < nsl; i += kPBlock) {
          const uint32_t c = cnt[i];
          if (c) {
            const uint32_t v = i < ns0 ? A.adj[r0.x + i] : A.adj[r1.x + (i - ns0)];
            atomicAdd(&A.counters[v], c);
            cnt[i] = 0u;
          }
        }
        lds_barrier();
      } else if (A.gen_hops > 1u) {
        // the first hops here: a tile's walkers share one or two sources, so
        // their successors are a few vertices (a few hundred after two hops)
        // whose rows stay in cache; the arrivals of these hops are aggregated
        // in the LDS table and added to the caller's counters, the last hop
        // is step_and_scatter's
        for (uint32_t hop = 0; hop + 1u < A.gen_hops; ++hop) {
          uint2 rd[kPer];
#pragma unroll
          for (uint32_t j = 0; j < kPer; ++j) rd[j] = valid[j] ? A.rowd[cur[j]] : make_uint2(0u, 0u);
#pragma unroll
          for (uint32_t j = 0; j < kPer; ++j) {
            valid[j] = valid[j] && rd[j].y != 0u;
            if (valid[j]) {
              const uint32_t k = __umulhi(sec[j] ^ fmix32(cur[j]), rd[j].y);
              const uint32_t ns = sec[j] * rd[j].y;
              if (ns == 0u && sec[j] != 0u) mark_bouquet(A.taint, wid[j]);  // hop + 1 < depth
              sec[j] = ns;
              cur[j] = A.adj[rd[j].x + k];
              count_arrival(L, hkey, hval, hlist, H, cur[j], A.counters);
            }
          }
        }
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) cur[j] = valid[j] ? cur[j] : kNone;
        step_and_scatter<false, kOutW>(L, A, A.gen_hops - 1u, nullptr, 0u, cur, sec, wid, pending,
                                [] {});
        // the barrier at the end of step_and_scatter follows every insert
        const uint32_t nh = L.hn;
        for (uint32_t i = tid; i < nh; i += kPBlock) {
          const uint32_t h = hlist[i];
          atomicAdd(&A.counters[hkey[h]], hval[h]);
          hkey[h] = kNone;
          hval[h] = 0u;
        }
        lds_barrier();
        if (tid == 0) L.hn = 0u;
        lds_barrier();
      } else {
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) cur[j] = valid[j] ? cur[j] : kNone;
        step_and_scatter<false, kOutW>(L, A, 0u, nullptr, 0u, cur, sec, wid, pending, [] {});
      }
    }
  } else {
    uint32_t* copy = A.copies + static_cast<uint64_t>(l) * A.n;
    for (uint32_t p = 0; p < A.P; ++p) {
      if (L.pxcd[p] != xc) continue;  // uniform
      const uint32_t nbp = A.bcount_in[p];
      if (l >= nbp) continue;  // uniform per block
      const uint32_t lo = L.plo[p];
      const uint32_t V = L.plo[p + 1] - lo;
      for (uint32_t i = tid; i < (kCnt16 ? (V + 1u) / 2u : V); i += kPBlock) cnt[i] = 0u;
      // adds the partition's counts into this workgroup's copy and clears them
      auto flush = [&]() {
        if constexpr (kCnt16) {
          for (uint32_t i = tid; i < (V + 1u) / 2u; i += kPBlock) {
            const uint32_t x = cnt[i];
            if (x & 0xFFFFu) copy[lo + 2u * i] += x & 0xFFFFu;
            if (x >> 16) copy[lo + 2u * i + 1u] += x >> 16;
            cnt[i] = 0u;
          }
        } else {
          for (uint32_t i = tid; i < V; i += kPBlock) {
            const uint32_t c = cnt[i];
            if (c) copy[lo + i] += c;
          }
        }
      };
      uint32_t since_flush = 0;
      if constexpr (kMode == kMid)
        for (uint32_t i = tid; i <= V; i += kPBlock) ridx[i] = A.idx[lo + i];
      const uint32_t base = A.pstart_in[p];
      const uint32_t mine = (nbp - l + NL - 1u) / NL;  // blocks l, l + NL, ... of p
      for (uint32_t b0 = 0; b0 < mine; b0 += kList) {
        const uint32_t nl = min(kList, mine - b0);
        lds_barrier();
        if (tid < nl) {
          const uint32_t blk = A.order_in[base + l + (b0 + tid) * NL];
          L.lblk[tid] = blk;
          L.lfill[tid] = A.bfill_in[blk];
        }
        lds_barrier();
        uint32_t cur[kPer], sec[kPer], wid[kPer];
        load_tile<kInW>(A, L.lblk[0], L.lfill[0], cur, sec, wid);
        for (uint32_t t = 0; t < nl; ++t) {
          // the next tile's records are in flight while this one is stepped
          uint32_t ncur[kPer], nsec[kPer], nwid[kPer];
          load_tile<kInW>(A, t + 1 < nl ? L.lblk[t + 1] : kNone,
                          t + 1 < nl ? L.lfill[t + 1] : 0u, ncur, nsec, nwid);
          if (kCnt16 && ++since_flush > kFlushTiles) {  // uniform
            lds_barrier();
            flush();
            lds_barrier();
            since_flush = 1;
          }
          auto arrive = [&] {
#pragma unroll
            for (uint32_t j = 0; j < kPer; ++j) {
              if (cur[j] != kNone) {
                const uint32_t o = cur[j] - lo;
                if constexpr (kCnt16) atomicAdd(&cnt[o >> 1], 1u << ((o & 1u) << 4));
                else atomicAdd(&cnt[o], 1u);
              }
            }
          };
          if constexpr (kMode == kMid)
            step_and_scatter<true, kOutW>(L, A, A.round, ridx, lo, cur, sec, wid, pending, arrive);
          else
            arrive();
#pragma unroll
          for (uint32_t j = 0; j < kPer; ++j) {
            cur[j] = ncur[j];
            sec[j] = nsec[j];
            wid[j] = nwid[j];
          }
        }
      }
      lds_barrier();
      flush();
    }
  }
  if constexpr (kMode != kLast)
    if (pending) copy_out<kOutW>(L, A);
  if constexpr (kMode != kLast) {
    lds_barrier();
    if (tid < A.P) {
      if (L.curblk[tid] != kNone) A.bfill_out[L.curblk[tid]] = L.curfill[tid];
      A.cwd[static_cast<uint64_t>(tid) * gridDim.x + w] = L.lcount[tid];
    }
    if (tid == 0) {
      A.reg_start[w] = L.rbase;
      A.reg_used[w] = min(L.used, L.rsize);
    }
  }
}

// Per partition d (one 64-thread block each): off[d][w] = blocks that
// workgroups before w wrote for d, bcount[d] = all of them.
__global__ void __launch_bounds__(64)
part_scan_kernel(const uint32_t* __restrict__ cwd, uint32_t grid, uint32_t* __restrict__ off,
                 uint32_t* __restrict__ bcount) {
  const uint32_t d = blockIdx.x, lane = threadIdx.x;
  const uint32_t* c = cwd + static_cast<uint64_t>(d) * grid;
  uint32_t* o = off + static_cast<uint64_t>(d) * grid;
  uint32_t run = 0;
  for (uint32_t w0 = 0; w0 < grid; w0 += 512u) {
    uint32_t x[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
      const uint32_t w = w0 + lane * 8u + j;
      x[j] = w < grid ? c[w] : 0u;
    }
    uint32_t s = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) s += x[j];
    uint32_t incl = s;
    for (int k = 1; k < 64; k <<= 1) {
      const uint32_t y = __shfl_up(incl, k);
      if (lane >= static_cast<uint32_t>(k)) incl += y;
    }
    uint32_t e = run + incl - s;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
      const uint32_t w = w0 + lane * 8u + j;
      if (w < grid) o[w] = e;
      e += x[j];
    }
    run += __shfl(incl, 63);
  }
  if (lane == 0) bcount[d] = run;
}

// Lists the blocks of the pool just written by partition:
// order[pstart[d] + off[d][w] + bslot[b]] = b for each block b of region w.
__global__ void __launch_bounds__(256)
part_order_kernel(const uint32_t* __restrict__ bcount, const uint32_t* __restrict__ off,
                  const uint32_t* __restrict__ bpart, const uint32_t* __restrict__ bslot,
                  const uint32_t* __restrict__ reg_start, const uint32_t* __restrict__ reg_used,
                  uint32_t grid, uint32_t P, uint32_t* __restrict__ order,
                  uint32_t* __restrict__ pstart) {
  __shared__ uint32_t pst[kMaxParts];
  __shared__ uint32_t wsum[4];
  const uint32_t tid = threadIdx.x;
  uint32_t v = tid < P ? bcount[tid] : 0u, incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o);
    if ((tid & 63u) >= static_cast<uint32_t>(o)) incl += y;
  }
  if ((tid & 63u) == 63u) wsum[tid >> 6] = incl;
  __syncthreads();
  uint32_t add = 0;
  for (uint32_t w = 0; w < (tid >> 6); ++w) add += wsum[w];
  pst[tid] = incl - v + add;
  __syncthreads();
  if (blockIdx.x == 0 && tid < P) pstart[tid] = pst[tid];
  for (uint32_t w = blockIdx.x; w < grid; w += gridDim.x) {
    const uint32_t b0 = reg_start[w], nb = reg_used[w];
    for (uint32_t i = tid; i < nb; i += 256u) {
      const uint32_t b = b0 + i;
      const uint32_t d = bpart[b];
      order[pst[d] + off[static_cast<uint64_t>(d) * grid + w] + bslot[b]] = b;
    }
  }
}

// out[e] = adj[e] | (partition of adj[e]) << 24; vertices below 2^24.
__global__ void __launch_bounds__(256)
part_tag_kernel(const uint32_t* __restrict__ adj, uint64_t m, const uint32_t* __restrict__ plo,
                uint32_t P, uint32_t* __restrict__ out) {
  __shared__ uint32_t s_plo[kMaxParts + 1];
  for (uint32_t i = threadIdx.x; i <= P; i += 256u) s_plo[i] = plo[i];
  __syncthreads();
  for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < m; e += gridDim.x * 256ull) {
    const uint32_t v = adj[e];
    uint32_t a = 0, b = P - 1u;  // largest p with plo[p] <= v
    while (a < b) {
      const uint32_t mid = (a + b + 1u) >> 1;
      if (s_plo[mid] <= v) a = mid;
      else b = mid - 1u;
    }
    out[e] = v | (a << 24);
  }
}

__global__ void part_fill_kernel(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = 0u;
}

// counters[v] += sum over copies
__global__ void __launch_bounds__(256)
part_reduce_kernel(const uint32_t* __restrict__ copies, uint32_t n_copies, uint64_t n,
                   uint32_t* __restrict__ counters) {
  for (uint64_t v = blockIdx.x * 256ull + threadIdx.x; v < n; v += gridDim.x * 256ull) {
    uint32_t s = 0;
    for (uint32_t c = 0; c < n_copies; ++c) s += copies[c * n + v];
    counters[v] += s;
  }
}

struct FixParams {
  const uint2* rowd;
  const uint32_t* adj;
  const uint64_t* wadj;  // packed {id | row start | degree} layout, or null
  uint32_t ib, sb;       // its id and row-start field widths
  const uint32_t* taint;
  uint32_t* counters;
  uint32_t n_words;
  uint32_t W;
  uint32_t step;
  uint32_t depth;
  uint32_t src_first;
  uint32_t src_stride;
};

// A walker's position: its vertex and that vertex's row (start, degree).
struct Pos {
  uint32_t v, start, deg;
};

__device__ __forceinline__ Pos pos_at(const FixParams& F, uint32_t v) {
  const uint2 rd = F.rowd[v];
  return Pos{v, rd.x, rd.y};
}

// The position after taking row entry k: one load with the packed layout,
// the neighbour and then its row otherwise.
template <bool kPacked>
__device__ __forceinline__ Pos step_to(const FixParams& F, const Pos& p, uint32_t k) {
  if constexpr (kPacked) {
    const uint64_t e = F.wadj[p.start + k];
    return Pos{static_cast<uint32_t>(e & ((1ull << F.ib) - 1ull)),
               static_cast<uint32_t>((e >> F.ib) & ((1ull << F.sb) - 1ull)),
               static_cast<uint32_t>(e >> (F.ib + F.sb))};
  } else {
    return pos_at(F, F.adj[p.start + k]);
  }
}

// Replays every marked bouquet, one 8-lane group per bouquet, a lane per
// walker, the group in lockstep: t* = first hop at whose start a live lane
// holds sec == 0 (a lane at a vertex without neighbours stops and leaves the
// test); from t* on each lane's independent visits are subtracted and its
// SplitMix32 visits added, the two walks interleaved (two chains of loads in
// flight).  A wave scans 64 bitmap words at a time and hands their marked
// bouquets to its eight groups.
template <bool kPacked>
__global__ void __launch_bounds__(256)
part_fixup_kernel(FixParams F) {
  const uint32_t lane = threadIdx.x & 63u, grp = lane >> 3, b = lane & 7u;
  const uint64_t gmask = 0xFFull << (grp * 8u);
  const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6;
  const uint32_t n_waves = (gridDim.x * 256u) >> 6;
  for (uint32_t base = wave * 64u; base < F.n_words; base += n_waves * 64u) {
    const uint32_t mine = base + lane < F.n_words ? F.taint[base + lane] : 0u;
    uint64_t nz = __ballot(mine != 0u);
    while (nz) {  // uniform per wave
      const uint32_t L0 = static_cast<uint32_t>(__builtin_ctzll(nz));
      nz &= nz - 1ull;
      uint32_t bits = __shfl(mine, static_cast<int>(L0));
      const uint32_t word = base + L0;
      while (bits) {
        // group g takes the g-th lowest marked bouquet of what is left
        uint32_t left = bits;
        for (uint32_t g = 0; g < grp && left; ++g) left &= left - 1u;
        const bool has = left != 0u;
        const uint32_t bq = has ? word * 32u + static_cast<uint32_t>(__builtin_ctz(left)) : 0u;
        for (uint32_t g = 0; g < 8u && bits; ++g) bits &= bits - 1u;
        const uint32_t wid0 = bq * 8u;
        const uint32_t sl = wid0 / F.W;
        const uint32_t init = (wid0 - sl * F.W + b + 1u) * F.step;
        const uint32_t src = F.src_first + sl * F.src_stride;
        // lockstep over the group's hops until one starts with a zero secret
        Pos p = has ? pos_at(F, src) : Pos{src, 0u, 0u};
        uint32_t sec = init;
        bool live = has;
        uint32_t ts = F.depth;
        for (uint32_t t = 0; t < F.depth; ++t) {
          const bool now = live && p.deg != 0u;
          // (each group's lanes take the same branch: the ballot is per group)
          if ((__ballot(now && sec == 0u) & gmask) != 0ull) {
            ts = t;
            break;
          }
          live = now;
          if (live) {
            const uint32_t k = __umulhi(sec ^ fmix32(p.v), p.deg);
            sec *= p.deg;
            p = step_to<kPacked>(F, p, k);
          }
        }
        if (ts < F.depth && live) {
          Pos p1 = p, p2 = p;
          uint32_t s1 = sec, st = init;
          bool a1 = true, a2 = true;
          for (uint32_t t = ts; t < F.depth && (a1 || a2); ++t) {
            a1 = a1 && p1.deg != 0u;
            a2 = a2 && p2.deg != 0u;
            if (a1) {
              const uint32_t k = __umulhi(s1 ^ fmix32(p1.v), p1.deg);
              s1 *= p1.deg;
              p1 = step_to<kPacked>(F, p1, k);
            }
            if (a2) p2 = step_to<kPacked>(F, p2, splitmix32(st) % p2.deg);
            if (a1) atomicAdd(&F.counters[p1.v], 0xFFFFFFFFu);
            if (a2) atomicAdd(&F.counters[p2.v], 1u);
          }
        }
      }
    }
  }
}

// Partition plan of a graph: contiguous vertex ranges, each with at most
// vmax vertices and about 1/want of the rows' and adjacency's bytes.
bool plan_parts(const std::vector<uint32_t>& deg, uint32_t vmax, uint32_t want,
                std::vector<uint32_t>& plo) {
  const uint64_t n = deg.size();
  double total = 0;
  for (uint64_t v = 0; v < n; ++v) total += 4.0 * deg[v] + 8.0;
  double target = total / std::max<uint32_t>(1, want);
  for (int tries = 0; tries < 64; ++tries) {
    plo.assign(1, 0u);
    double acc = 0;
    uint32_t V = 0;
    for (uint64_t v = 0; v < n; ++v) {
      const double c = 4.0 * deg[v] + 8.0;
      if (V > 0 && (V >= vmax || acc + c > target)) {
        plo.push_back(static_cast<uint32_t>(v));
        acc = 0;
        V = 0;
        if (plo.size() > kMaxParts) break;
      }
      acc += c;
      ++V;
    }
    if (plo.size() <= kMaxParts) {
      plo.push_back(static_cast<uint32_t>(n));
      return true;
    }
    target *= 1.25;
  }
  return false;
}

// XCD of each partition (the residue mod 8 of the workgroups serving it).
// Balanced: heaviest first to the XCD with the least expected records so
// far, a partition weighing its degree sum (a walk visits a vertex at a rate
// proportional to its degree once it has mixed); each XCD takes its
// partitions in increasing order.  Otherwise p % 8.
std::vector<uint8_t> assign_xcds(const std::vector<uint32_t>& deg,
                                 const std::vector<uint32_t>& plo, bool balance) {
  const uint32_t P = static_cast<uint32_t>(plo.size() - 1);
  std::vector<uint8_t> x(kMaxParts, 0);
  for (uint32_t p = 0; p < P; ++p) x[p] = static_cast<uint8_t>(p & 7u);
  if (!balance) return x;
  std::vector<double> w(P, 0.0);
  for (uint32_t p = 0; p < P; ++p)
    for (uint32_t v = plo[p]; v < plo[p + 1]; ++v) w[p] += deg[v];
  std::vector<uint32_t> order(P);
  for (uint32_t p = 0; p < P; ++p) order[p] = p;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return w[a] > w[b]; });
  double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t p : order) {
    uint32_t best = 0;
    for (uint32_t k = 1; k < 8; ++k)
      if (load[k] < load[best]) best = k;
    x[p] = static_cast<uint8_t>(best);
    load[best] += w[p];
  }
  return x;
}

struct PartShape {
  uint32_t P = 0, vbig = 0, grid = 0, NL = 0, pool_cap = 0;
  size_t lds = 0;
  uint64_t n_chunks = 0, chunk_src = 0;
};

bool part_env_on(const tgtp_dgraph* dg) {
  // TGTP_WALK_PART: -1 auto (graphs whose rows and adjacency exceed
  // TGTP_WALK_PART_MIN_MIB), 0 off, 1 required (an error when not taken)
  const int mode = env_int("TGTP_WALK_PART", -1);
  if (mode == 0 || dg->m == 0 || dg->n > 0xFFFFFFFFull || dg->m > 0xFFFFFFFFull) return false;
  const double footprint = 4.0 * static_cast<double>(dg->m) + 8.0 * static_cast<double>(dg->n);
  const double min_mib = env_int("TGTP_WALK_PART_MIN_MIB", 64);
  return mode > 0 || footprint >= min_mib * 1048576.0;
}

// Launch shape and chunking of a pass; false: the engine does not take it.
bool part_shape(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint64_t n_src, uint32_t n_walks,
                PartShape* S) {
  if (dg->part_plo.size() < 2 || n_src == 0 || n_walks == 0) return false;
  S->P = static_cast<uint32_t>(dg->part_plo.size() - 1);
  S->vbig = 0;
  for (uint32_t p = 0; p < S->P; ++p)
    S->vbig = std::max(S->vbig, dg->part_plo[p + 1] - dg->part_plo[p]);
  // every workgroup resident, a multiple of 8 (the XCDs)
  S->lds = sizeof(PartLds) +
           ((kCnt16 ? (S->vbig + 1ull) / 2 : S->vbig) + S->vbig + 1ull) * sizeof(uint32_t);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void*>(&part_round_kernel<kMid, 3, 3>), kPBlock, S->lds) !=
          hipSuccess ||
      per_cu < 1)
    return false;
  per_cu = std::min(per_cu, std::max(1, env_int("TGTP_WALK_PART_PER_CU", 2)));
  S->grid = std::max<uint32_t>(8u, static_cast<uint32_t>(ctx->num_cus * per_cu) / 8u * 8u);
  S->NL = S->grid / 8u;
  // chunk: sources whose walkers fit the record pools
  // Fewer, larger chunks save rounds, and each round costs a fixed share
  // (a partial block per partition and workgroup, the tail of the launch):
  // the pools may take up to 96 GiB, where the 32-bit record index ends, but
  // at most 3/4 of the memory free beside what they hold already (the free
  // memory is not queried once they hold the cap).
  const uint64_t pool_cap_bytes =
      static_cast<uint64_t>(std::max(64, env_int("TGTP_WALK_PART_MIB", 98304))) << 20;
  // a round's regions hold its input blocks plus P per workgroup, and its
  // input holds at most the walkers' full blocks plus P partial blocks per
  // workgroup of the round before
  const uint64_t slack_blocks = 2ull * S->grid * S->P;
  const uint64_t blk_bytes = static_cast<uint64_t>(kBR) * 12u + 4u * 4u;  // records + tables
  const uint64_t have = ctx->bufs[kBufPartRecA].bytes + ctx->bufs[kBufPartRecB].bytes;
  // blocks of a pool holding the whole pass in one chunk, and the bytes of
  // both such pools
  const uint64_t blocks_all = (n_src * n_walks + kBR - 1) / kBR + slack_blocks;
  uint64_t budget = pool_cap_bytes;
  if (have >= pool_cap_bytes) {
    budget = pool_cap_bytes;
  } else if (n_src * n_walks <= 0xFFFFFFF8ull && blocks_all + 1 <= kMaxPoolBlocks &&
             have >= 2ull * (blocks_all + 1) * kBR * 12u) {
    budget = 2ull * (blocks_all + 1) * blk_bytes;  // the pools already hold it: one chunk
  } else {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
    budget = std::min<uint64_t>(pool_cap_bytes, std::max<uint64_t>(have, (free_b + have) / 4 * 3));
  }
  const uint64_t pool_blocks_max = std::min<uint64_t>(kMaxPoolBlocks, budget / 2 / blk_bytes);
  if (pool_blocks_max <= slack_blocks + 1) return false;
  const uint64_t walkers_max =
      std::min<uint64_t>((pool_blocks_max - slack_blocks) * kBR, 0xFFFFFFF8ull);
  uint64_t chunk_src = walkers_max / n_walks;
  if (chunk_src == 0) return false;
  if (const int c = env_int("TGTP_WALK_PART_CHUNK", 0); c > 0)
    chunk_src = std::min<uint64_t>(chunk_src, static_cast<uint64_t>(c));
  chunk_src = std::min<uint64_t>(chunk_src, n_src);
  S->n_chunks = (n_src + chunk_src - 1) / chunk_src;
  S->chunk_src = (n_src + S->n_chunks - 1) / S->n_chunks;
  const uint64_t walkers = S->chunk_src * n_walks;
  const uint64_t cap = (walkers + kBR - 1) / kBR + slack_blocks;
  if (cap + 1 > kMaxPoolBlocks) return false;
  S->pool_cap = static_cast<uint32_t>(cap);
  return true;
}

}  // namespace

int saba_part_prepare(tgtp_ctx* ctx, tgtp_dgraph* dg, uint64_t n_src, uint32_t depth,
                      uint32_t n_walks, bool* use) {
  *use = false;
  if (depth == 0 || !part_env_on(dg)) return TGTP_OK;
  const uint64_t n = dg->n;
  const uint32_t vmax = static_cast<uint32_t>(std::max(64, env_int("TGTP_WALK_PART_VMAX", kCnt16 ? 10240 : 8192)));
  const uint32_t want = static_cast<uint32_t>(std::max(1, env_int("TGTP_WALK_PARTS", 128)));
  const bool required = env_int("TGTP_WALK_PART", -1) > 0;
  auto refuse = [&](const char* why) {
    return required ? set_error(ctx, TGTP_E_UNSUPPORTED,
                                std::string("walk: partitioned engine required but ") + why)
                    : TGTP_OK;
  };
  // TGTP_WALK_PART_BALANCE=0: partition p to the workgroups b with b % 8 == p % 8
  const uint32_t balance = env_int("TGTP_WALK_PART_BALANCE", 1) != 0 ? 1u : 0u;
  if (dg->part_plo.empty() || dg->part_vmax != vmax || dg->part_want != want ||
      dg->part_balance != balance) {
    dg->part_plo.clear();
    dg->part_tagged = false;
    std::vector<uint2> rd(n);
    TGTP_HIP_CHECK(ctx, hipMemcpy(rd.data(), dg->rowd, n * sizeof(uint2), hipMemcpyDeviceToHost));
    std::vector<uint32_t> deg(n);
    for (uint64_t v = 0; v < n; ++v) deg[v] = rd[v].y;
    std::vector<uint32_t> plo;
    if (!plan_parts(deg, vmax, want, plo)) return refuse("the graph needs more vertex ranges than sort bins");
    const uint32_t P = static_cast<uint32_t>(plo.size() - 1);
    uint32_t tshift = 0;
    while (((n + (1ull << tshift) - 1) >> tshift) > kTab) ++tshift;
    std::vector<uint16_t> ptab(kTab + 2, static_cast<uint16_t>(P - 1));
    uint32_t p = 0;
    for (uint64_t t = 0; t < kTab && (t << tshift) < n; ++t) {
      while (plo[p + 1] <= (t << tshift)) ++p;
      ptab[t] = static_cast<uint16_t>(p);
    }
    if (!dg->part_dev) {
      TGTP_HIP_CHECK(ctx, hipMalloc(&dg->part_dev, (kMaxParts + 1) * sizeof(uint32_t) +
                                                       (kTab + 2) * sizeof(uint16_t) + kMaxParts));
    }
    const std::vector<uint8_t> pxcd = assign_xcds(deg, plo, balance != 0);
    TGTP_HIP_CHECK(ctx, hipMemcpy(static_cast<uint8_t*>(dg->part_dev) +
                                      (kMaxParts + 1) * sizeof(uint32_t) +
                                      (kTab + 2) * sizeof(uint16_t),
                                  pxcd.data(), kMaxParts, hipMemcpyHostToDevice));
    TGTP_HIP_CHECK(ctx, hipMemcpy(dg->part_dev, plo.data(), plo.size() * sizeof(uint32_t),
                                  hipMemcpyHostToDevice));
    TGTP_HIP_CHECK(ctx, hipMemcpy(static_cast<uint8_t*>(dg->part_dev) +
                                      (kMaxParts + 1) * sizeof(uint32_t),
                                  ptab.data(), ptab.size() * sizeof(uint16_t),
                                  hipMemcpyHostToDevice));
    dg->part_plo = plo;
    dg->part_vmax = vmax;
    dg->part_want = want;
    dg->part_tshift = tshift;
    dg->part_balance = balance;
    dg->part_tagged = false;
  }
  // the neighbour gathers of the rounds read the partition with the vertex
  // (TGTP_WALK_PART_TAGS=0: looked up in the boundaries instead)
  if (n <= (1ull << 24) && env_int("TGTP_WALK_PART_TAGS", 1) != 0 && !dg->part_tagged) {
    if (!dg->part_adj)
      TGTP_HIP_CHECK(ctx, hipMalloc(&dg->part_adj, std::max<uint64_t>(dg->m, 1) * sizeof(uint32_t)));
    part_tag_kernel<<<static_cast<uint32_t>(std::min<uint64_t>(8192, (dg->m + 255) / 256 + 1)), 256,
                      0, ctx->stream>>>(dg->adj, dg->m, static_cast<const uint32_t*>(dg->part_dev),
                                        static_cast<uint32_t>(dg->part_plo.size() - 1),
                                        dg->part_adj);
    TGTP_HIP_CHECK(ctx, hipGetLastError());
    dg->part_tagged = true;
  }
  PartShape S;
  *use = part_shape(ctx, dg, n_src, n_walks, &S);
  return *use ? TGTP_OK : refuse("the pass does not fit its record pools");
}

// Walks of sources [src_begin, src_begin + n_src) with the partitioned
// engine, adding into d_counters (cleared by the caller); saba_part_prepare
// has reported that the pass takes it.
int launch_saba_walks_part(tgtp_ctx* ctx, tgtp_dgraph* dg, uint32_t src_begin, uint64_t n_src,
                           uint32_t depth, uint32_t n_walks, uint32_t* d_counters) {
  PartShape S;
  if (!part_shape(ctx, dg, n_src, n_walks, &S))
    return set_error(ctx, TGTP_E_UNSUPPORTED, "walk: partitioned engine not prepared");
  const uint64_t n = dg->n;
  const uint32_t P = S.P, grid = S.grid, NL = S.NL, pool_cap = S.pool_cap;
  const size_t lds = S.lds;
  const uint64_t n_chunks = S.n_chunks;
  const uint64_t walkers = S.chunk_src * n_walks;

  // buffers: two record pools with their block tables, the per-round
  // region tables, counter copies, bouquet bitmap
  const uint64_t rec_bytes = (static_cast<uint64_t>(pool_cap) + 1u) * kBR * 12u;
  // per pool: bfill, bpart, bslot, order [pool_cap] + bcount, pstart [kMaxParts]
  const uint64_t meta_words = 4ull * pool_cap + 2ull * kMaxParts;
  // shared: cwd, off [kMaxParts * grid] + reg_start, reg_used [grid] + overflow
  const uint64_t region_words = 2ull * kMaxParts * grid + 2ull * grid + 64;
  const uint64_t bq_words = (walkers / 8 + 31) / 32;
  auto* recA = static_cast<uint32_t*>(ctx_buffer(ctx, kBufPartRecA, rec_bytes));
  auto* recB = static_cast<uint32_t*>(ctx_buffer(ctx, kBufPartRecB, rec_bytes));
  auto* meta = static_cast<uint32_t*>(
      ctx_buffer(ctx, kBufPartMeta, (2 * meta_words + region_words) * sizeof(uint32_t)));
  auto* copies = static_cast<uint32_t*>(ctx_buffer(ctx, kBufPartCopies, NL * n * 4));
  auto* taint = static_cast<uint32_t*>(ctx_buffer(ctx, kBufPartTaint, bq_words * 4));
  if (!recA || !recB || !meta || !copies || !taint)
    return set_error(ctx, TGTP_E_NOMEM, "walk: partition record pools");
  struct Pool {
    uint32_t *rec, *bfill, *bpart, *bslot, *order, *bcount, *pstart;
  } pool[2];
  for (int i = 0; i < 2; ++i) {
    uint32_t* m = meta + i * meta_words;
    pool[i].rec = i == 0 ? recA : recB;
    pool[i].bfill = m;
    pool[i].bpart = m + pool_cap;
    pool[i].bslot = m + 2ull * pool_cap;
    pool[i].order = m + 3ull * pool_cap;
    pool[i].bcount = m + 4ull * pool_cap;
    pool[i].pstart = pool[i].bcount + kMaxParts;
  }
  uint32_t* cwd = meta + 2 * meta_words;
  uint32_t* off = cwd + static_cast<uint64_t>(kMaxParts) * grid;
  uint32_t* reg_start = off + static_cast<uint64_t>(kMaxParts) * grid;
  uint32_t* reg_used = reg_start + grid;
  uint32_t* overflow = reg_used + grid;

  hipStream_t s = ctx->stream;
  auto fill = [&](uint32_t* p, uint64_t words) {
    part_fill_kernel<<<static_cast<uint32_t>(std::min<uint64_t>(4096, (words + 255) / 256 + 1)), 256, 0,
                       s>>>(p, words);
  };
  fill(copies, NL * n);
  fill(overflow, 1);

  PartRound A{};
  A.rowd = dg->rowd;
  A.idx = dg->idx;
  A.adj = dg->adj;
  A.vbig = S.vbig;
  A.plo = static_cast<const uint32_t*>(dg->part_dev);
  A.ptab = reinterpret_cast<const uint16_t*>(static_cast<const uint8_t*>(dg->part_dev) +
                                             (kMaxParts + 1) * sizeof(uint32_t));
  A.pxcd = static_cast<const uint8_t*>(dg->part_dev) + (kMaxParts + 1) * sizeof(uint32_t) +
           (kTab + 2) * sizeof(uint16_t);
  A.tadj = dg->part_tagged && env_int("TGTP_WALK_PART_TAGS", 1) != 0 ? dg->part_adj : nullptr;
  A.n = n;
  A.P = P;
  A.tshift = dg->part_tshift;
  A.depth = depth;
  A.W = n_walks;
  A.step = 0xFFFFFFFFu / n_walks;
  A.src_stride = static_cast<uint32_t>(n_chunks);
  A.pool_cap = pool_cap;
  A.counters = d_counters;
  A.gen_hops = std::min<uint32_t>(depth, static_cast<uint32_t>(
                                           std::max(1, env_int("TGTP_WALK_PART_GEN_HOPS", 2))));
  A.gen_slots = A.gen_hops == 2u && n_walks >= kBR &&
                2ull * dg->max_deg <= (kCnt16 ? (S.vbig + 1ull) / 2 : S.vbig) + S.vbig + 1 &&
                env_int("TGTP_WALK_PART_GEN_SLOTS", 1) != 0;
  A.cwd = cwd;
  A.reg_start = reg_start;
  A.reg_used = reg_used;
  A.overflow = overflow;
  A.copies = copies;
  A.taint = taint;
  FixParams F{};
  F.rowd = dg->rowd;
  F.adj = dg->adj;
  {
    // the replay's hops read the packed layout when it can be built (one
    // load per hop instead of two)
    bool packed = false;
    int prc = ensure_walk_adj(ctx, dg, &packed);
    if (prc) return prc;
    F.wadj = packed ? dg->walk_adj : nullptr;
    F.ib = dg->id_bits;
    F.sb = dg->start_bits;
  }
  F.taint = taint;
  F.counters = d_counters;
  F.W = n_walks;
  F.step = A.step;
  F.depth = depth;
  F.src_stride = static_cast<uint32_t>(n_chunks);

  for (uint64_t c = 0; c < n_chunks; ++c) {
    // chunk c takes every n_chunks-th source from src_begin + c
    const uint64_t c_n = (n_src - c + n_chunks - 1) / n_chunks;
    const uint32_t nw = static_cast<uint32_t>(c_n * n_walks);
    const uint64_t words = (static_cast<uint64_t>(nw) / 8 + 31) / 32;
    A.src_first = static_cast<uint32_t>(src_begin + c);
    A.n_walkers = nw;
    fill(taint, words);
    int cur = 0;
    for (uint32_t r = 0; r <= depth; ++r) {
      if (r > 0 && r < A.gen_hops) continue;  // taken by the generation round
      const Pool& in = pool[cur ^ 1];
      const Pool& out = pool[cur];
      A.round = r;
      A.rec_in = in.rec;
      A.order_in = in.order;
      A.pstart_in = in.pstart;
      A.bcount_in = in.bcount;
      A.bfill_in = in.bfill;
      A.rec_out = out.rec;
      A.bfill_out = out.bfill;
      A.bpart_out = out.bpart;
      A.bslot_out = out.bslot;
      // words per record into the round that takes hop h: 3, but 2 into the
      // round that takes the last hop (no walker id) and 1 into the last
      // round (the vertex)
      auto words_into = [&](uint32_t h) { return h == depth ? 1 : h + 1u == depth ? 2 : 3; };
      if (r == 0) {
        const int ow = words_into(A.gen_hops);
        if (ow == 1) part_round_kernel<kGen, 0, 1><<<grid, kPBlock, lds, s>>>(A);
        else if (ow == 2) part_round_kernel<kGen, 0, 2><<<grid, kPBlock, lds, s>>>(A);
        else part_round_kernel<kGen, 0, 3><<<grid, kPBlock, lds, s>>>(A);
      } else if (r < depth) {
        const int ow = words_into(r + 1u);  // 1: r is the last hop (its input: 2)
        if (ow == 1) part_round_kernel<kMid, 2, 1><<<grid, kPBlock, lds, s>>>(A);
        else if (ow == 2) part_round_kernel<kMid, 3, 2><<<grid, kPBlock, lds, s>>>(A);
        else part_round_kernel<kMid, 3, 3><<<grid, kPBlock, lds, s>>>(A);
      } else {
        part_round_kernel<kLast, 1, 0><<<grid, kPBlock, lds, s>>>(A);
      }
      TGTP_HIP_CHECK(ctx, hipGetLastError());
      if (r < depth) {
        // list the blocks just written by partition for the next round
        part_scan_kernel<<<P, 64, 0, s>>>(cwd, grid, off, out.bcount);
        part_order_kernel<<<256, 256, 0, s>>>(out.bcount, off, out.bpart, out.bslot, reg_start,
                                              reg_used, grid, P, out.order, out.pstart);
        TGTP_HIP_CHECK(ctx, hipGetLastError());
        cur ^= 1;
      }
    }
    F.src_first = A.src_first;
    F.n_words = static_cast<uint32_t>(words);
    const uint32_t fix_grid = static_cast<uint32_t>(std::min<uint64_t>(2048, (words + 255) / 256));
    if (F.wadj)
      part_fixup_kernel<true><<<fix_grid, 256, 0, s>>>(F);
    else
      part_fixup_kernel<false><<<fix_grid, 256, 0, s>>>(F);
    TGTP_HIP_CHECK(ctx, hipGetLastError());
  }
  part_reduce_kernel<<<static_cast<uint32_t>(std::min<uint64_t>(4096, (n + 255) / 256)), 256, 0,
                       s>>>(copies, NL, n, d_counters);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  uint32_t ovf = 0;
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(&ovf, overflow, 4, hipMemcpyDeviceToHost, s));
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if (ovf) return set_error(ctx, TGTP_E_HIP, "walk: partition record pool overflow");
  return TGTP_OK;
}
