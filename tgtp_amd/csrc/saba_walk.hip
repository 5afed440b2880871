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
// its neighbour gathers fall in one adjacency row.  Work is a grid-stride
// loop over bouquets with an incremental (source, bouquet) divmod, so the
// loop body holds no integer division; every block is resident at once.
//
// Memory: a hop is a chain of dependent random loads, served mostly by the
// Infinity Cache (the CSR of a 2^20-vertex, 20M-entry graph is ~90 MB).  When
// the field widths fit, the walk reads a packed layout built once per graph:
// walk_adj[p] = {neighbour id | neighbour row start | neighbour degree} in 64
// bits, so one random 8-byte load per hop yields everything the next hop
// needs (the plain layout needs the neighbour id, then its row descriptor).
// Counters of vertices [0, hot) are staged in LDS per workgroup and flushed
// once; visits to other vertices are device-scope atomics, which this part
// performs memory-side (one 32-byte sector each).
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
  uint32_t id_bits;     // packed layout: bits of the vertex id field
  uint32_t start_bits;  // packed layout: bits of the row-start field
};

__device__ __forceinline__ void count_visit(uint32_t v, uint32_t hot, uint32_t* hot_cnt,
                                            uint32_t* counters) {
  if (v < hot) {
    atomicAdd(&hot_cnt[v], 1u);
  } else {
    __hip_atomic_fetch_add(&counters[v], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <bool kPacked>
__global__ void __launch_bounds__(kWalkBlock)
saba_walk_kernel(const uint2* __restrict__ rowd, const uint32_t* __restrict__ adj,
                 const uint64_t* __restrict__ walk_adj, uint32_t* __restrict__ counters,
                 WalkParams P) {
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
  const uint64_t id_mask = (1ull << P.id_bits) - 1ull;
  const uint64_t start_mask = (1ull << P.start_bits) - 1ull;
  const uint32_t deg_shift = P.id_bits + P.start_bits;

  for (; base < P.total_bq; base += P.stride_bq) {
    const bool active = base + grp < P.total_bq;
    uint32_t cur = P.src_begin + s_off;
    uint32_t sec = (j * 8u + b + 1u) * P.step;
    uint32_t st = sec;
    bool fb = false;
    uint32_t start = 0u, deg = 0u;
    if (active) {
      const uint2 r = rowd[cur];
      start = r.x;
      deg = r.y;
    }
    bool live = active && deg != 0u;  // degree-0 sources are skipped
    for (uint32_t hop = 0; hop < P.depth; ++hop) {
      // a zero-degree vertex is reachable only in a non-symmetric CSR: the
      // walker stops there instead of reading past its row
      live = live && deg != 0u;
      const uint64_t z = __ballot(live && !fb && sec == 0u);
      fb = fb || (((z >> (grp * 8u)) & 0xFFull) != 0ull);
      if (live) {
        uint32_t k;
        if (!fb) {
          k = __umulhi(sec ^ fmix32(cur), deg);
          sec *= deg;
        } else {
          k = splitmix32(st) % deg;
        }
        if constexpr (kPacked) {
          const uint64_t e = walk_adj[start + k];
          cur = static_cast<uint32_t>(e & id_mask);
          start = static_cast<uint32_t>((e >> P.id_bits) & start_mask);
          deg = static_cast<uint32_t>(e >> deg_shift);
        } else {
          cur = adj[start + k];
          if (hop + 1 < P.depth) {
            const uint2 r = rowd[cur];
            start = r.x;
            deg = r.y;
          }
        }
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

__global__ void build_walk_adj_kernel(const uint32_t* __restrict__ idx,
                                      const uint32_t* __restrict__ adj, uint64_t m,
                                      uint32_t id_bits, uint32_t start_bits,
                                      uint64_t* __restrict__ walk_adj) {
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < m;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t v = adj[p];
    const uint64_t s = idx[v], d = idx[v + 1] - idx[v];
    walk_adj[p] = static_cast<uint64_t>(v) | (s << id_bits) | (d << (id_bits + start_bits));
  }
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t value) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = value;
}

// sums[0] += sum(c), sums[1] += sum((v+1) * c[v]); the caller clears sums
// with fill_u64_kernel first.
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

uint32_t bits_for(uint64_t x) {  // bits needed to hold values 0..x
  uint32_t b = 1;
  while (b < 64 && (x >> b) != 0) ++b;
  return b;
}

// Builds the packed walk layout once per graph when its three fields fit in
// 64 bits; *packed reports whether the walk can use it.
int ensure_walk_adj(tgtp_ctx* ctx, tgtp_dgraph* dg, bool* packed) {
  *packed = false;
  if (env_int("TGTP_WALK_PACKED", 1) == 0 || dg->m == 0) return TGTP_OK;
  const uint32_t ib = bits_for(dg->n - 1), sb = bits_for(dg->m), db = bits_for(dg->max_deg);
  if (ib + sb + db > 64) return TGTP_OK;
  if (!dg->walk_adj) {
    if (hipMalloc(&dg->walk_adj, dg->m * sizeof(uint64_t)) != hipSuccess) {
      dg->walk_adj = nullptr;  // not fatal: the plain layout still works
      return TGTP_OK;
    }
    dg->id_bits = ib;
    dg->start_bits = sb;
    const uint64_t blocks = std::min<uint64_t>(8192, (dg->m + 255) / 256);
    build_walk_adj_kernel<<<static_cast<uint32_t>(blocks), 256, 0, ctx->stream>>>(
        dg->idx, dg->adj, dg->m, ib, sb, dg->walk_adj);
    TGTP_HIP_CHECK(ctx, hipGetLastError());
  }
  *packed = true;
  return TGTP_OK;
}

}  // namespace

int launch_saba_walks(tgtp_ctx* ctx, const tgtp_dgraph* cdg, uint32_t src_begin,
                      uint32_t src_end, uint32_t depth, uint32_t n_walks,
                      uint32_t* d_counters, unsigned long long* d_sums, float* walk_ms) {
  auto* dg = const_cast<tgtp_dgraph*>(cdg);
  hipStream_t s = ctx->stream;
  const uint64_t n = dg->n;
  bool packed = false;
  int rc = ensure_walk_adj(ctx, dg, &packed);
  if (rc) return rc;
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
    P.id_bits = packed ? dg->id_bits : 0;
    P.start_bits = packed ? dg->start_bits : 0;
    // LDS staging of the lowest vertex ids (the hubs of the Chung-Lu
    // generator; correct for any graph, only the hit rate changes).  32K
    // counters = 128 KiB: one 256-thread block per CU.
    const uint32_t hot_max = static_cast<uint32_t>(env_int("TGTP_WALK_HOT", 32768));
    P.hot = static_cast<uint32_t>(std::min<uint64_t>(n, hot_max));
    const size_t lds = static_cast<size_t>(P.hot) * sizeof(uint32_t);
    // every block resident at once: the grid-stride loop gives each block the
    // same share, so a second batch of blocks would double the tail
    int per_cu = 0;
    const void* fn = packed ? reinterpret_cast<const void*>(&saba_walk_kernel<true>)
                            : reinterpret_cast<const void*>(&saba_walk_kernel<false>);
    TGTP_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWalkBlock, lds));
    per_cu = std::max(1, std::min(per_cu, env_int("TGTP_WALK_BLOCKS_PER_CU", 64)));
    const uint64_t tasks = (total_bq + 7) / 8;
    const uint64_t blocks_needed = (tasks + (kWalkBlock / 64) - 1) / (kWalkBlock / 64);
    const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks_needed,
                                               static_cast<uint64_t>(ctx->num_cus) * per_cu));
    P.stride_bq = blocks * (kWalkBlock / 64) * 8ull;
    P.stride_src = static_cast<uint32_t>(P.stride_bq / bps);
    P.stride_rem = static_cast<uint32_t>(P.stride_bq % bps);
    const dim3 grid(static_cast<uint32_t>(blocks));
    if (packed)
      saba_walk_kernel<true><<<grid, kWalkBlock, lds, s>>>(dg->rowd, dg->adj, dg->walk_adj,
                                                          d_counters, P);
    else
      saba_walk_kernel<false><<<grid, kWalkBlock, lds, s>>>(dg->rowd, dg->adj, nullptr,
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
