// mt_walk.hip -- the reference's MT19937 comparator walks on the GPU:
//   regular  rw_experiment.cpp:34-47   walk i, hop h draws output i*depth + h
//   simd     rw_experiment_simd.cpp:44-82
//            bouquets of 8 walks; hop h of bouquet b draws the 8 outputs
//            b*8*depth + 8h + (0..7), one per lane
// of the source's own std::mt19937 stream, seeded with 42 + source
// (rw_experiment.cpp:103); the next vertex is adj[row + draw % degree], and
// every reached vertex is counted.  Counters are uint32 per vertex (the
// reference keeps per-thread float counters).
//
// 256-thread workgroups, each looping over sources.  The generator is MT19937 as
// published (Matsumoto & Nishimura, ACM TOMACS 8(1), 1998; the parameters
// std::mt19937 fixes): its 624-word state sits in LDS, seeded by one lane
// (the seeding recurrence is sequential), and each twist runs in four
// phases across the lanes -- words [0, 227) read only old words, [227, 454)
// and [454, 623) read words the previous phase wrote, word 623 reads new
// words 0 and 396.  Tempered outputs go to a 4096-word ring in LDS; a batch
// of up to 256 walks (one per thread) runs once the ring holds its whole
// range of the stream, so each thread reads its draws by index.  Visits of
// the 4096 lowest ids (the hubs of the seeded Chung-Lu graphs, whose
// counters every walk hits) count in LDS and are added to HBM once per
// workgroup; other visits are device atomics.
#include <algorithm>

#include "common.hpp"

namespace {

constexpr uint32_t kMtN = 624;
constexpr uint32_t kMtM = 397;
constexpr uint32_t kMtUpper = 0x80000000u;
constexpr uint32_t kMtLower = 0x7fffffffu;
constexpr uint32_t kMtA = 0x9908b0dfu;
constexpr uint32_t kRing = 4096;    // tempered outputs kept per source
constexpr uint32_t kMtBlock = 256;  // threads (walks in flight) per source
constexpr uint32_t kMtHot = 4096;   // vertices counted in LDS

struct MtWalkArgs {
  const uint2* rowd;  // {row start, degree}
  const uint32_t* adj;
  const uint64_t* wadj;  // packed {id | row start | degree} adjacency, or null
  uint32_t ib, sb;       // its id and row-start field widths
  uint32_t* counts;
  uint32_t n;
  uint32_t src_begin, src_end;
  uint32_t depth, n_walks;
  uint32_t walks_per_batch;  // <= kMtBlock; a multiple of 8 for the simd variant
  int32_t variant;           // 0 regular, 1 simd
};

// std::mt19937::seed: x[0] = s, x[i] = 1812433253 (x[i-1] ^ x[i-1] >> 30) + i.
__device__ void mt_seed(uint32_t* mt, uint32_t seed) {
  if (threadIdx.x == 0) {
    uint32_t x = seed;
    mt[0] = x;
    for (uint32_t i = 1; i < kMtN; ++i) {
      x = 1812433253u * (x ^ (x >> 30)) + i;
      mt[i] = x;
    }
  }
  __syncthreads();
}

// New word i from (old or already new) words i, i+1 and i+397 (mod 624).
__device__ __forceinline__ uint32_t mt_word(const uint32_t* mt, uint32_t i) {
  const uint32_t y = (mt[i] & kMtUpper) | (mt[i + 1 == kMtN ? 0 : i + 1] & kMtLower);
  const uint32_t j = i + kMtM < kMtN ? i + kMtM : i + kMtM - kMtN;
  return mt[j] ^ (y >> 1) ^ ((y & 1u) ? kMtA : 0u);
}

// Words [lo, hi) in chunks of kMtBlock: each chunk's words are read before
// any is written (word i + 1 of the chunk's last thread belongs to the next
// chunk).
__device__ __forceinline__ void mt_phase(uint32_t* mt, uint32_t lo, uint32_t hi) {
  for (uint32_t c = lo; c < hi; c += kMtBlock) {
    const uint32_t i = c + threadIdx.x;
    uint32_t v = 0;
    if (i < hi) v = mt_word(mt, i);
    __syncthreads();
    if (i < hi) mt[i] = v;
    __syncthreads();
  }
}

__device__ void mt_twist(uint32_t* mt) {
  mt_phase(mt, 0, kMtN - kMtM);                  // reads old words only
  mt_phase(mt, kMtN - kMtM, 2 * (kMtN - kMtM));  // word i - 227 is new
  mt_phase(mt, 2 * (kMtN - kMtM), kMtN - 1);     // word i - 227 is new
  mt_phase(mt, kMtN - 1, kMtN);                  // words 0 and 396 are new
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

// kPacked: a hop reads one packed 8-byte entry (the next vertex, its row
// start and degree) instead of the neighbour id and then its row.
template <bool kPacked>
__global__ void __launch_bounds__(kMtBlock) mt_walk_kernel(MtWalkArgs a) {
  __shared__ uint32_t mt[kMtN];
  __shared__ uint32_t ring[kRing];
  __shared__ uint32_t hot[kMtHot];
  const uint32_t lane = threadIdx.x;
  for (uint32_t v = lane; v < kMtHot; v += kMtBlock) hot[v] = 0;
  __syncthreads();
  const uint64_t depth = a.depth;
  for (uint32_t s = a.src_begin + blockIdx.x; s < a.src_end; s += gridDim.x) {
    if (a.rowd[s].y == 0) continue;  // isolated source: skipped (rw_experiment.cpp:99-100)
    mt_seed(mt, 42u + s);
    uint64_t produced = 0;           // stream outputs tempered so far
    for (uint32_t w0 = 0; w0 < a.n_walks; w0 += a.walks_per_batch) {
      const uint32_t nw = min(a.walks_per_batch, a.n_walks - w0);
      // the batch draws outputs [w0 * depth, (w0 + nw) * depth) in both variants
      const uint64_t o_end = static_cast<uint64_t>(w0 + nw) * depth;
      while (produced < o_end) {
        mt_twist(mt);
        for (uint32_t i = lane; i < kMtN; i += kMtBlock)
          ring[(produced + i) % kRing] = mt_temper(mt[i]);
        __syncthreads();
        produced += kMtN;
      }
      if (lane < nw) {
        const uint32_t w = w0 + lane;
        uint32_t cur = s;
        uint2 d = a.rowd[s];  // {row start, degree} of cur
        for (uint32_t h = 0; h < a.depth; ++h) {
          const uint64_t o = a.variant == 0
                                 ? static_cast<uint64_t>(w) * depth + h
                                 : static_cast<uint64_t>(w / 8u) * 8u * depth + 8u * h + (w % 8u);
          const uint32_t r = ring[o % kRing];
          // a vertex without neighbours (only in a non-symmetric CSR; the
          // reference divides by zero there): the walker stops, and since
          // each hop's draw index is fixed, no other walk moves in the stream
          if (d.y == 0) break;
          if constexpr (kPacked) {
            const uint64_t e = a.wadj[d.x + r % d.y];
            cur = static_cast<uint32_t>(e & ((1ull << a.ib) - 1ull));
            d = make_uint2(static_cast<uint32_t>((e >> a.ib) & ((1ull << a.sb) - 1ull)),
                           static_cast<uint32_t>(e >> (a.ib + a.sb)));
          } else {
            cur = a.adj[d.x + r % d.y];
            d = a.rowd[cur];
          }
          if (cur < kMtHot)
            atomicAdd(&hot[cur], 1u);
          else
            atomicAdd(&a.counts[cur], 1u);
        }
      }
      __syncthreads();  // the batch's draws are read before the ring moves on
    }
  }
  for (uint32_t v = lane; v < kMtHot && v < a.n; v += kMtBlock)
    if (hot[v]) atomicAdd(&a.counts[v], hot[v]);
}

// One std::mt19937 stream seeded `seed`: outputs [0, count) in order, one
// workgroup (the generator is sequential; the twist is spread over its lanes).
__global__ void __launch_bounds__(kMtBlock) mt_stream_kernel(uint32_t seed, uint64_t count,
                                                           uint32_t* __restrict__ out) {
  __shared__ uint32_t mt[kMtN];
  mt_seed(mt, seed);
  for (uint64_t base = 0; base < count; base += kMtN) {
    mt_twist(mt);
    for (uint32_t i = threadIdx.x; i < kMtN; i += kMtBlock)
      if (base + i < count) out[base + i] = mt_temper(mt[i]);
    __syncthreads();
  }
}

}  // namespace

int launch_mt_stream(tgtp_ctx* ctx, uint32_t seed, uint64_t count, uint32_t* d_out) {
  if (count == 0) return TGTP_OK;
  mt_stream_kernel<<<1, kMtBlock, 0, ctx->stream>>>(seed, count, d_out);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}

int launch_mt_walks(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin, uint32_t src_end,
                    uint32_t depth, uint32_t n_walks, int32_t variant, uint32_t* d_counters,
                    unsigned long long* d_sums, float* walk_ms, bool may_alloc) {
  hipStream_t s = ctx->stream;
  // a batch's range of the stream plus one twist must fit the ring
  uint32_t wpb = depth ? std::min<uint32_t>(kMtBlock, (kRing - kMtN) / depth) : kMtBlock;
  if (variant == 1) wpb -= wpb % 8u;
  if (wpb == 0)
    return set_error(ctx, TGTP_E_UNSUPPORTED,
                     variant == 1 ? "depth above 434 for the simd variant"
                                  : "depth above 3472 for the regular variant");
  int rc;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], s));
  if ((rc = launch_walk_counters_clear(ctx, d_counters, dg->n, d_sums))) return rc;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], s));
  if (src_end > src_begin && depth > 0 && n_walks > 0 && dg->m > 0) {
    MtWalkArgs a{};
    a.rowd = dg->rowd;
    a.adj = dg->adj;
    // the packed layout when it can be built here (or is there already: the
    // entry point that allocates nothing only reads it)
    bool packed = dg->walk_adj != nullptr;
    if (may_alloc && (rc = ensure_walk_adj(ctx, const_cast<tgtp_dgraph*>(dg), &packed))) return rc;
    a.wadj = packed ? dg->walk_adj : nullptr;
    a.ib = dg->id_bits;
    a.sb = dg->start_bits;
    a.counts = d_counters;
    a.n = static_cast<uint32_t>(dg->n);
    a.src_begin = src_begin;
    a.src_end = src_end;
    a.depth = depth;
    a.n_walks = n_walks;
    a.walks_per_batch = wpb;
    a.variant = variant;
    // workgroups loop over sources: enough of them to fill the GPU, few
    // enough that the per-workgroup flush of the LDS counters stays small
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&mt_walk_kernel<true>), kMtBlock, 0) !=
        hipSuccess)
      per_cu = 1;
    const uint32_t groups = std::min<uint32_t>(
        src_end - src_begin, static_cast<uint32_t>(std::max(1, per_cu)) * ctx->num_cus);
    if (a.wadj)
      mt_walk_kernel<true><<<groups, kMtBlock, 0, s>>>(a);
    else
      mt_walk_kernel<false><<<groups, kMtBlock, 0, s>>>(a);
    TGTP_HIP_CHECK(ctx, hipGetLastError());
  }
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[2], s));
  if ((rc = launch_walk_counter_sums(ctx, d_counters, dg->n, d_sums))) return rc;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[3], s));
  TGTP_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev[3]));
  if (walk_ms) TGTP_HIP_CHECK(ctx, hipEventElapsedTime(walk_ms, ctx->ev[1], ctx->ev[2]));
  return TGTP_OK;
}
