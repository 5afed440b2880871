// saba_walk.hip -- SABA 1-way walk engine for gfx950.
//
// Semantics (rw_experiment_saba.cpp:65-145): for a source s and W walks,
// walk i = 8*j + b (bouquet j, lane b) starts at s with grid secret
// sec = (8*j + b + 1) * floor((2^32-1)/W) mod 2^32.  Each hop, if any lane
// of the bouquet holds sec == 0 the whole bouquet switches for good to the
// fallback k = splitmix32(state) % deg (state seeded with the initial
// secret); otherwise k = hi32((sec ^ fmix32(cur)) * deg) and sec *= deg.
// Then cur = adj[idx[cur] + k] and cur's counter is incremented.
//
// Mapping: one 8-lane bouquet per 8 consecutive lanes, eight bouquets per
// 64-lane wavefront.  The bouquet-wide zero test is one wave ballot whose
// 8-bit group is tested per bouquet.  Consecutive bouquets of one source
// share a wave, so the first hop's row descriptor is one broadcast load and
// its neighbour gathers fall in one adjacency row.  Counters of vertices
// [0, hot) are staged in LDS (per workgroup) and flushed with one global
// atomic per nonzero entry at the end; the rest go straight to L2 atomics.
// Work is a grid-stride loop over bouquets with an incremental
// (source, bouquet) divmod, so the loop body holds no integer division.
#include "common.hpp"

#include <algorithm>
#include <cstdlib>

namespace {

constexpr int kWalkBlock = 256;

struct WalkParams {
  uint64_t total_bq;    // bouquets in this launch
  uint64_t stride_bq;   // bouquets advanced per grid-stride step (8 * waves)
  uint32_t src_begin;
  uint32_t bps;         // bouquets per source = W / 8
  uint32_t stride_src;  // stride_bq / bps
  uint32_t stride_rem;  // stride_bq % bps
  uint32_t depth;
  uint32_t step;        // floor((2^32-1) / W)
  uint32_t hot;         // vertices counted in LDS
};

__device__ __forceinline__ void count_visit(uint32_t v, uint32_t hot, uint32_t* hot_cnt,
                                            uint32_t* counters) {
  if (v < hot) {
    atomicAdd(&hot_cnt[v], 1u);
  } else {
    __hip_atomic_fetch_add(&counters[v], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void __launch_bounds__(kWalkBlock)
saba_walk_kernel(const uint2* __restrict__ rowd, const uint32_t* __restrict__ adj,
                 uint32_t* __restrict__ counters, WalkParams P) {
  extern __shared__ uint32_t hot_cnt[];
  for (uint32_t i = threadIdx.x; i < P.hot; i += kWalkBlock) hot_cnt[i] = 0;
  __syncthreads();

  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t grp = lane >> 3;  // bouquet slot in the wave
  const uint32_t b = lane & 7u;    // lane in the bouquet
  const uint64_t wave = static_cast<uint64_t>(blockIdx.x) * (kWalkBlock / 64) + (threadIdx.x >> 6);
  uint64_t base = wave * 8u;
  const uint64_t bq0 = base + grp;
  uint32_t s_off = static_cast<uint32_t>(bq0 / P.bps);
  uint32_t j = static_cast<uint32_t>(bq0 % P.bps);

  for (; base < P.total_bq; base += P.stride_bq) {
    const bool active = base + grp < P.total_bq;
    uint32_t cur = P.src_begin + s_off;
    uint32_t sec = (j * 8u + b + 1u) * P.step;
    uint32_t st = sec;
    bool fb = false;
    uint2 rd = make_uint2(0u, 0u);
    if (active) rd = rowd[cur];
    bool live = active && rd.y != 0u;  // degree-0 sources are skipped
    for (uint32_t hop = 0; hop < P.depth; ++hop) {
      // a zero-degree vertex is reachable only in a non-symmetric CSR: the
      // walker stops there instead of reading past its row
      live = live && rd.y != 0u;
      const uint64_t z = __ballot(live && !fb && sec == 0u);
      fb = fb || (((z >> (grp * 8u)) & 0xFFull) != 0ull);
      if (live) {
        const uint32_t d = rd.y;
        uint32_t k;
        if (!fb) {
          k = __umulhi(sec ^ fmix32(cur), d);
          sec *= d;
        } else {
          k = splitmix32(st) % d;
        }
        cur = adj[rd.x + k];
        if (hop + 1 < P.depth) rd = rowd[cur];
        count_visit(cur, P.hot, hot_cnt, counters);
      }
    }
    s_off += P.stride_src;
    j += P.stride_rem;
    if (j >= P.bps) {
      j -= P.bps;
      ++s_off;
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < P.hot; i += kWalkBlock) {
    const uint32_t c = hot_cnt[i];
    if (c) __hip_atomic_fetch_add(&counters[i], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t value) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = value;
}

// sums[0] += sum(c), sums[1] += sum((v+1) * c[v]); sums zeroed by the
// first block's thread 0 is not possible across blocks, so the caller
// clears them with fill_u64_kernel first.
__global__ void __launch_bounds__(256)
counter_sums_kernel(const uint32_t* __restrict__ c, uint64_t n, unsigned long long* sums) {
  unsigned long long s0 = 0, s1 = 0;
  for (uint64_t v = blockIdx.x * 256ull + threadIdx.x; v < n; v += gridDim.x * 256ull) {
    const unsigned long long x = c[v];
    s0 += x;
    s1 += (v + 1) * x;
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off);
    s1 += __shfl_xor(s1, off);
  }
  if ((threadIdx.x & 63u) == 0) {
    atomicAdd(&sums[0], s0);
    atomicAdd(&sums[1], s1);
  }
}

__global__ void fill_u64_kernel(unsigned long long* p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

int env_int(const char* name, int dflt) {
  const char* s = std::getenv(name);
  return s && *s ? std::atoi(s) : dflt;
}

}  // namespace

int launch_saba_walks(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin,
                      uint32_t src_end, uint32_t depth, uint32_t n_walks,
                      uint32_t* d_counters, unsigned long long* d_sums, float* walk_ms) {
  hipStream_t s = ctx->stream;
  const uint64_t n = dg->n;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], s));
  fill_u32_kernel<<<std::min<uint64_t>(4096, (n + 255) / 256 + 1), 256, 0, s>>>(d_counters, n, 0u);
  fill_u64_kernel<<<1, 64, 0, s>>>(d_sums, 2);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], s));

  const uint64_t n_src = src_end > src_begin ? src_end - src_begin : 0;
  const uint32_t bps = n_walks / 8u;
  const uint64_t total_bq = n_src * bps;
  if (total_bq > 0 && depth > 0) {
    WalkParams P{};
    P.total_bq = total_bq;
    P.src_begin = src_begin;
    P.bps = bps;
    P.depth = depth;
    P.step = 0xFFFFFFFFu / n_walks;
    // LDS staging of the lowest vertex ids (the hubs of the Chung-Lu
    // generator; correct for any graph, only the hit rate changes).
    const uint32_t hot_max = static_cast<uint32_t>(env_int("TGTP_WALK_HOT", 8192));
    P.hot = static_cast<uint32_t>(std::min<uint64_t>(n, hot_max));
    const int per_cu = env_int("TGTP_WALK_BLOCKS_PER_CU", 8);
    const uint64_t waves_needed = (total_bq + 7) / 8;
    const uint64_t blocks_needed = (waves_needed + (kWalkBlock / 64) - 1) / (kWalkBlock / 64);
    const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks_needed,
                                               static_cast<uint64_t>(ctx->num_cus) * per_cu));
    P.stride_bq = blocks * (kWalkBlock / 64) * 8ull;
    P.stride_src = static_cast<uint32_t>(P.stride_bq / bps);
    P.stride_rem = static_cast<uint32_t>(P.stride_bq % bps);
    const size_t lds = static_cast<size_t>(P.hot) * sizeof(uint32_t);
    saba_walk_kernel<<<static_cast<uint32_t>(blocks), kWalkBlock, lds, s>>>(dg->rowd, dg->adj,
                                                                           d_counters, P);
    TGTP_HIP_CHECK(ctx, hipGetLastError());
  }
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[2], s));
  counter_sums_kernel<<<std::min<uint64_t>(1024, (n + 255) / 256 + 1), 256, 0, s>>>(d_counters, n,
                                                                                  d_sums);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[3], s));
  TGTP_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev[3]));
  if (walk_ms) TGTP_HIP_CHECK(ctx, hipEventElapsedTime(walk_ms, ctx->ev[1], ctx->ev[2]));
  return TGTP_OK;
}
