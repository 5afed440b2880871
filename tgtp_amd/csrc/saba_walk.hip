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
// needs (the plain layout needs the neighbour id, then its row descriptor);
// the hop loop is software-pipelined, hop h+1's gather in flight while hop
// h's visit is counted.
//
// Counting: counters of vertices [0, hot) live in LDS per workgroup and are
// flushed once.  A visit to another ("cold") vertex would be a device-scope
// atomic, which this part performs memory-side (one 32-byte sector each);
// instead it is staged in LDS per vertex-range bucket and written out 32 at a
// time as a 128-byte record run into the wave's own region of HBM, and
// bucket_count_kernel then counts each bucket's records in an LDS histogram.
// Sources run in strided chunks so that a chunk's records fit the regions; a
// region that fills up falls back to the atomic, so every count is exact.
#include "common.hpp"

#include <algorithm>
#include <cstdlib>

namespace {

constexpr int kWalkBlock = 256;
constexpr int kWalkWaves = kWalkBlock / 64;
// Bucketed cold counting: vertices [hot, n) are split into at most
// kMaxBuckets ranges of 2^bshift; each wave stages its cold visits per range
// in LDS and writes them out kFlush at a time (128 B) into its own region of
// HBM, and bucket_count_kernel then counts every range with LDS atomics.
constexpr int kMaxBuckets = 16;
constexpr uint32_t kFlush = 32;
constexpr uint32_t kStageCap = kFlush + 64;  // < kFlush before a hop, +64 lanes in it
constexpr uint32_t kHistBins = 32768;         // vertices per LDS histogram pass
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct WalkParams {
  uint64_t total_bq;    // bouquets in this launch
  uint64_t stride_bq;   // bouquets advanced per grid-stride step (8 * waves)
  uint32_t src_begin;
  uint32_t src_stride;  // sources of this launch: src_begin + i * src_stride
  uint32_t bps;         // bouquets per source = W / 8
  uint32_t stride_src;  // stride_bq / bps
  uint32_t stride_rem;  // stride_bq % bps
  uint32_t depth;
  uint32_t step;        // floor((2^32-1) / W)
  uint32_t hot;         // vertices counted in LDS
  uint32_t id_bits;     // packed layout: bits of the vertex id field
  uint32_t start_bits;  // packed layout: bits of the row-start field
  // bucketed mode
  uint32_t n_buckets;
  uint32_t bshift;
  uint32_t rbase[kMaxBuckets + 1];  // region layout of one wave (records)
  uint32_t* rec;                    // [waves][rbase[n_buckets]]
  uint32_t* fill;                   // [waves][n_buckets] records written
};

struct StageLds {
  uint32_t cnt[kWalkWaves][kMaxBuckets];
  uint32_t off[kWalkWaves][kMaxBuckets];
  uint32_t rec[kWalkWaves][kMaxBuckets][kStageCap];
};

// Writes a staged visit to record slot `idx` of the wave's region of bucket
// `bk`, or counts it directly when the region is full.
__device__ __forceinline__ void emit_record(const WalkParams& P, uint32_t* counters,
                                            uint64_t region, uint32_t bk, uint32_t idx,
                                            uint32_t v) {
  if (idx < P.rbase[bk + 1] - P.rbase[bk])
    P.rec[region + P.rbase[bk] + idx] = v;
  else
    __hip_atomic_fetch_add(&counters[v], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool kPacked, bool kBucket>
__global__ void __launch_bounds__(kWalkBlock)
saba_walk_kernel(const uint2* __restrict__ rowd, const uint32_t* __restrict__ adj,
                 const uint64_t* __restrict__ walk_adj, uint32_t* __restrict__ counters,
                 WalkParams P) {
  extern __shared__ uint32_t lds[];
  uint32_t* hot_cnt = lds;
  StageLds* stage = reinterpret_cast<StageLds*>(lds + P.hot);
  for (uint32_t i = threadIdx.x; i < P.hot; i += kWalkBlock) hot_cnt[i] = 0;
  if constexpr (kBucket) {
    for (uint32_t i = threadIdx.x; i < kWalkWaves * kMaxBuckets; i += kWalkBlock) {
      (&stage->cnt[0][0])[i] = 0;
      (&stage->off[0][0])[i] = 0;
    }
  }
  __syncthreads();

  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t grp = lane >> 3;  // bouquet slot in the wave
  const uint32_t b = lane & 7u;    // lane in the bouquet
  const uint64_t wave = static_cast<uint64_t>(blockIdx.x) * kWalkWaves + wv;
  const uint64_t region = wave * P.rbase[P.n_buckets];
  uint64_t base = wave * 8u;
  const uint64_t bq0 = base + grp;
  uint32_t s_off = static_cast<uint32_t>(bq0 / P.bps);
  uint32_t j = static_cast<uint32_t>(bq0 % P.bps);
  const uint64_t id_mask = (1ull << P.id_bits) - 1ull;
  const uint64_t start_mask = (1ull << P.start_bits) - 1ull;
  const uint32_t deg_shift = P.id_bits + P.start_bits;

  for (; base < P.total_bq; base += P.stride_bq) {  // uniform per wave
    const bool active = base + grp < P.total_bq;
    uint32_t cur = P.src_begin + s_off * P.src_stride;
    uint32_t sec = (j * 8u + b + 1u) * P.step;
    uint32_t st = sec;
    bool fb = false;
    const uint2 r0 = rowd[active ? cur : 0u];
    uint32_t start = active ? r0.x : 0u, deg = active ? r0.y : 0u;
    bool live = active && deg != 0u;  // degree-0 sources are skipped
    // Counts one visit (kNone: none).  Converged: every lane of the wave calls it.
    auto count_visit = [&](uint32_t vis) {
      if (vis != kNone && vis < P.hot) atomicAdd(&hot_cnt[vis], 1u);
      if constexpr (!kBucket) {
        if (vis != kNone && vis >= P.hot)
          __hip_atomic_fetch_add(&counters[vis], 1u, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
      } else {
        const bool cold = vis != kNone && vis >= P.hot;
        uint32_t bk = 0u, pos = 0u;
        if (cold) {
          bk = (vis - P.hot) >> P.bshift;
          pos = atomicAdd(&stage->cnt[wv][bk], 1u);
          stage->rec[wv][bk][pos] = vis;
        }
        // flush every bucket that reached kFlush in this hop
        uint64_t need = __ballot(cold && pos == kFlush - 1u);
        while (need) {
          const int l = __builtin_ctzll(need);
          need &= need - 1ull;
          const uint32_t bb = __shfl(bk, l);
          uint32_t cnt = stage->cnt[wv][bb];
          uint32_t off = stage->off[wv][bb];
          while (cnt >= kFlush) {
            if (lane < kFlush) emit_record(P, counters, region, bb, off + lane, stage->rec[wv][bb][lane]);
            const uint32_t rem = cnt - kFlush;
            const uint32_t t = lane < rem ? stage->rec[wv][bb][kFlush + lane] : 0u;
            __builtin_amdgcn_wave_barrier();
            if (lane < rem) stage->rec[wv][bb][lane] = t;
            cnt = rem;
            off += kFlush;
          }
          __builtin_amdgcn_wave_barrier();
          if (lane == 0) {
            stage->cnt[wv][bb] = cnt;
            stage->off[wv][bb] = off;
          }
          __builtin_amdgcn_wave_barrier();
        }
      }
    };
    // Picks the next neighbour slot of a live lane (the reference's hash
    // step, or the bouquet's SplitMix32 fallback once any lane's secret is 0).
    auto step = [&](bool lv) -> uint32_t {
      const uint64_t z = __ballot(lv && !fb && sec == 0u);
      fb = fb || (((z >> (grp * 8u)) & 0xFFull) != 0ull);
      uint32_t k = 0u;
      if (lv) {
        if (!fb) {
          k = __umulhi(sec ^ fmix32(cur), deg);
          sec *= deg;
        } else {
          k = splitmix32(st) % deg;
        }
      }
      return k;
    };
    if constexpr (kPacked) {
      // Software-pipelined: hop h+1's gather is issued before hop h's visit
      // is counted, so the LDS atomics and staging of a visit overlap the
      // next gather instead of preceding it.  A lane not walking reads entry
      // 0 (m > 0 when the kernel runs).
      const uint32_t k0 = step(live);  // every lane: the ballot inside must be converged
      uint64_t e = walk_adj[live ? start + k0 : 0u];
      for (uint32_t hop = 0; hop < P.depth; ++hop) {
        uint32_t vis = kNone;
        if (live) {
          cur = static_cast<uint32_t>(e & id_mask);
          start = static_cast<uint32_t>((e >> P.id_bits) & start_mask);
          deg = static_cast<uint32_t>(e >> deg_shift);
          vis = cur;
        }
        if (hop + 1 < P.depth) {
          // a zero-degree vertex is reachable only in a non-symmetric CSR:
          // the walker stops there instead of reading past its row
          live = live && deg != 0u;
          const uint32_t k = step(live);
          e = walk_adj[live ? start + k : 0u];
        }
        count_visit(vis);
      }
    } else {
      for (uint32_t hop = 0; hop < P.depth; ++hop) {
        live = live && deg != 0u;
        const uint32_t k = step(live);
        uint32_t vis = kNone;
        if (live) {
          cur = adj[start + k];
          if (hop + 1 < P.depth) {
            const uint2 r = rowd[cur];
            start = r.x;
            deg = r.y;
          }
          vis = cur;
        }
        count_visit(vis);
      }
    }
    s_off += P.stride_src;
    j += P.stride_rem;
    if (j >= P.bps) {
      j -= P.bps;
      ++s_off;
    }
  }
  if constexpr (kBucket) {
    // partial stages of this wave, then the per-bucket record counts
    for (uint32_t bb = 0; bb < P.n_buckets; ++bb) {
      const uint32_t cnt = stage->cnt[wv][bb];
      const uint32_t off = stage->off[wv][bb];
      if (lane < cnt) emit_record(P, counters, region, bb, off + lane, stage->rec[wv][bb][lane]);
      if (lane == 0)
        P.fill[wave * P.n_buckets + bb] = min(off + cnt, P.rbase[bb + 1] - P.rbase[bb]);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < P.hot; i += kWalkBlock) {
    const uint32_t c = hot_cnt[i];
    if (c) __hip_atomic_fetch_add(&counters[i], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

struct HistParams {
  uint64_t n;
  uint32_t hot;
  uint32_t n_buckets;
  uint32_t bshift;
  uint32_t passes;       // LDS passes per bucket: 2^bshift / kHistBins (>= 1)
  uint32_t slices;       // wave slices per (bucket, pass)
  uint32_t waves;
  uint32_t rtot;         // records per wave region
  uint32_t rbase[kMaxBuckets + 1];
};

constexpr int kCountBlock = 1024;  // 16 waves share one LDS histogram

// blockIdx.x = (bucket * passes + pass) * slices + slice: counts the records
// of waves [slice * waves / slices, ...) that fall in the pass's vertex range.
// Records are read 4 at a time (regions start 128-byte aligned), two loads
// in flight per thread, into one LDS histogram per workgroup.
__global__ void __launch_bounds__(kCountBlock)
bucket_count_kernel(const uint32_t* __restrict__ rec, const uint32_t* __restrict__ fill,
                    uint32_t* __restrict__ counters, HistParams Q) {
  __shared__ uint32_t hist[kHistBins];
  const uint32_t slice = blockIdx.x % Q.slices;
  const uint32_t bp = blockIdx.x / Q.slices;
  const uint32_t bk = bp / Q.passes, pass = bp % Q.passes;
  const uint64_t lo64 = Q.hot + (static_cast<uint64_t>(bk) << Q.bshift) +
                        static_cast<uint64_t>(pass) * kHistBins;
  if (lo64 >= Q.n) return;  // uniform per block
  const uint32_t lo = static_cast<uint32_t>(lo64);
  const uint32_t span = static_cast<uint32_t>(min<uint64_t>(kHistBins, Q.n - lo64));
  for (uint32_t i = threadIdx.x; i < kHistBins; i += kCountBlock) hist[i] = 0;
  __syncthreads();
  auto add = [&](uint32_t v) {
    const uint32_t d = v - lo;
    if (d < span) atomicAdd(&hist[d], 1u);
  };
  const uint32_t w0 = static_cast<uint32_t>(static_cast<uint64_t>(slice) * Q.waves / Q.slices);
  const uint32_t w1 = static_cast<uint32_t>(static_cast<uint64_t>(slice + 1) * Q.waves / Q.slices);
  for (uint32_t w = w0; w < w1; ++w) {
    const uint32_t cnt = fill[static_cast<uint64_t>(w) * Q.n_buckets + bk];
    const uint32_t* r = rec + static_cast<uint64_t>(w) * Q.rtot + Q.rbase[bk];
    const uint4* r4 = reinterpret_cast<const uint4*>(r);
    const uint32_t n4 = cnt / 4;
    uint32_t i = threadIdx.x;
    for (; i + kCountBlock < n4; i += 2 * kCountBlock) {
      const uint4 a = r4[i];
      const uint4 c = r4[i + kCountBlock];
      add(a.x); add(a.y); add(a.z); add(a.w);
      add(c.x); add(c.y); add(c.z); add(c.w);
    }
    if (i < n4) {
      const uint4 a = r4[i];
      add(a.x); add(a.y); add(a.z); add(a.w);
    }
    if (threadIdx.x < cnt - 4 * n4) add(r[4 * n4 + threadIdx.x]);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < span; i += kCountBlock) {
    const uint32_t c = hist[i];
    if (c) __hip_atomic_fetch_add(&counters[lo + i], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

}  // namespace

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

namespace {

const void* walk_kernel_ptr(bool packed, bool bucket) {
  if (packed)
    return bucket ? reinterpret_cast<const void*>(&saba_walk_kernel<true, true>)
                  : reinterpret_cast<const void*>(&saba_walk_kernel<true, false>);
  return bucket ? reinterpret_cast<const void*>(&saba_walk_kernel<false, true>)
                : reinterpret_cast<const void*>(&saba_walk_kernel<false, false>);
}

// Expected share of visits per cold bucket, from the degree mass of its
// vertex range (a walk's visits follow the degrees once it has mixed); only
// sizes the record regions, a region that fills up falls back to atomics.
// Depends on the graph alone: kept on the dgraph for the next call.
std::vector<double> bucket_mass(tgtp_ctx* ctx, tgtp_dgraph* dg, uint32_t hot, uint32_t bshift,
                                uint32_t nb) {
  if (!dg->walk_mass.empty() && dg->walk_mass.size() == nb && dg->walk_mass_hot == hot &&
      dg->walk_mass_bshift == bshift)
    return dg->walk_mass;
  std::vector<double> mass(nb, 0.0);
  std::vector<uint2> rd(dg->n);
  if (hipMemcpy(rd.data(), dg->rowd, dg->n * sizeof(uint2), hipMemcpyDeviceToHost) != hipSuccess)
    return {};
  double tot = 0.0;
  for (uint64_t v = 0; v < dg->n; ++v) {
    tot += rd[v].y;
    if (v >= hot) mass[(v - hot) >> bshift] += rd[v].y;
  }
  for (auto& x : mass) x = tot > 0 ? x / tot : 0.0;
  (void)ctx;
  dg->walk_mass = mass;
  dg->walk_mass_hot = hot;
  dg->walk_mass_bshift = bshift;
  return mass;
}

}  // namespace

int launch_walk_counters_clear(tgtp_ctx* ctx, uint32_t* d_counters, uint64_t n,
                               unsigned long long* d_sums) {
  fill_u32_kernel<<<std::min<uint64_t>(4096, (n + 255) / 256 + 1), 256, 0, ctx->stream>>>(
      d_counters, n, 0u);
  fill_u64_kernel<<<1, 64, 0, ctx->stream>>>(d_sums, 2);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}

int launch_walk_counter_sums(tgtp_ctx* ctx, const uint32_t* d_counters, uint64_t n,
                             unsigned long long* d_sums) {
  counter_sums_kernel<<<std::min<uint64_t>(1024, (n + 255) / 256 + 1), 256, 0, ctx->stream>>>(
      d_counters, n, d_sums);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}

int launch_saba_walks(tgtp_ctx* ctx, const tgtp_dgraph* cdg, uint32_t src_begin,
                      uint32_t src_end, uint32_t depth, uint32_t n_walks,
                      uint32_t* d_counters, unsigned long long* d_sums, float* walk_ms) {
  auto* dg = const_cast<tgtp_dgraph*>(cdg);
  hipStream_t s = ctx->stream;
  const uint64_t n = dg->n;
  const uint64_t n_src = src_end > src_begin ? src_end - src_begin : 0;
  const uint32_t bps = n_walks / 8u;
  const bool any = n_src > 0 && bps > 0 && depth > 0 && dg->m > 0;  // m == 0: no vertex can step
  // large graphs: the partitioned engine (saba_part.hip), planned before the
  // timed window like the packed layout
  bool part = false;
  int rc = any ? saba_part_prepare(ctx, dg, n_src, depth, n_walks, &part) : TGTP_OK;
  if (rc) return rc;
  bool packed = false;
  if (!part && (rc = ensure_walk_adj(ctx, dg, &packed))) return rc;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], s));
  if ((rc = launch_walk_counters_clear(ctx, d_counters, n, d_sums))) return rc;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], s));

  if (any && part) {
    if ((rc = launch_saba_walks_part(ctx, dg, src_begin, n_src, depth, n_walks, d_counters)))
      return rc;
  } else if (any) {
    WalkParams P{};
    P.src_begin = src_begin;
    P.bps = bps;
    P.depth = depth;
    P.step = 0xFFFFFFFFu / n_walks;
    P.id_bits = packed ? dg->id_bits : 0;
    P.start_bits = packed ? dg->start_bits : 0;
    // LDS counters of the lowest vertex ids (the hubs of the Chung-Lu
    // generator; correct for any graph, only the hit rate changes).  With
    // bucketed cold counting the default is the largest table that leaves
    // two workgroups per CU (more walks in flight beat a higher LDS hit
    // rate: 14,208 counters measured 8% faster than 32,768 at n = 2^20);
    // with direct atomics a cold visit is expensive and 32,768 wins.
    const bool want_bucket = env_int("TGTP_WALK_BUCKETS", 1) != 0;
    uint32_t hot_dflt = 32768;
    if (want_bucket) {
      int lds_cu = 0;
      if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                ctx->device) != hipSuccess || lds_cu <= 0)
        lds_cu = 65536;
      const int64_t room = (static_cast<int64_t>(lds_cu) / 2 -
                            static_cast<int64_t>(sizeof(StageLds))) / 4;
      hot_dflt = static_cast<uint32_t>(std::max<int64_t>(0, room) / 64 * 64);
    }
    const uint32_t hot_max = static_cast<uint32_t>(env_int("TGTP_WALK_HOT", static_cast<int>(hot_dflt)));
    P.hot = static_cast<uint32_t>(std::min<uint64_t>(n, hot_max));
    // cold vertices in at most kMaxBuckets ranges of 2^bshift >= kHistBins
    bool bucket = want_bucket && n > P.hot;
    uint32_t nb = 0, passes = 1;
    std::vector<double> mass;
    if (bucket) {
      P.bshift = 15;
      while (((n - P.hot + (1ull << P.bshift) - 1) >> P.bshift) > kMaxBuckets) ++P.bshift;
      nb = static_cast<uint32_t>((n - P.hot + (1ull << P.bshift) - 1) >> P.bshift);
      passes = static_cast<uint32_t>((1ull << P.bshift) / kHistBins);
      // each record is read once per LDS pass: beyond 4 passes the counting
      // pass costs more than the atomics it replaces
      if (passes > 4) bucket = false;
      if (bucket) {
        mass = bucket_mass(ctx, dg, P.hot, P.bshift, nb);
        if (mass.empty()) return set_error(ctx, TGTP_E_HIP, "walk: rowd copy failed");
      }
    }
    P.n_buckets = bucket ? nb : 0;
    const size_t lds = static_cast<size_t>(P.hot) * sizeof(uint32_t) + (bucket ? sizeof(StageLds) : 0);
    const void* fn = walk_kernel_ptr(packed, bucket);
    // every block resident at once: the grid-stride loop gives each block the
    // same share, so a second batch of blocks would double the tail
    int per_cu = 0;
    TGTP_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWalkBlock, lds));
    if (per_cu < 1) return set_error(ctx, TGTP_E_UNSUPPORTED, "walk: LDS request does not fit a CU");
    per_cu = std::min(per_cu, env_int("TGTP_WALK_BLOCKS_PER_CU", 64));
    const uint64_t resident = static_cast<uint64_t>(ctx->num_cus) * per_cu;

    // sources per launch: all of them, or (bucketed) as many as the record
    // regions hold at the expected load
    uint64_t chunk_src = n_src;
    uint64_t waves = 0;
    uint32_t* rec = nullptr;
    uint32_t* fill = nullptr;
    HistParams Q{};
    if (bucket) {
      waves = resident * kWalkWaves;
      const uint64_t have = ctx->bufs[kBufWalkRec].bytes;
      const uint64_t cap = static_cast<uint64_t>(env_int("TGTP_WALK_REC_MIB", 16384)) << 20;
      uint64_t budget = cap;
      if (have < cap) {  // the record regions have not reached their cap yet
        size_t free_b = 0, total_b = 0;
        TGTP_HIP_CHECK(ctx, hipMemGetInfo(&free_b, &total_b));
        budget = std::min<uint64_t>(cap, std::max<uint64_t>(have, (free_b + have) / 4));
      }
      double cold_share = 0.0;
      for (double x : mass) cold_share += x;
      const double per_src = static_cast<double>(n_walks) * depth;
      // records per wave of bucket k for c sources: c * per_src * mass[k] / waves,
      // with 25% margin (TGTP_WALK_REC_PCT) plus two flushes of slack
      const double margin = std::max(1, env_int("TGTP_WALK_REC_PCT", 125)) / 100.0;
      const double rec_words = static_cast<double>(budget) / 4.0 / static_cast<double>(waves) -
                               2.0 * kFlush * nb - kFlush * nb;
      if (rec_words <= 0 || cold_share <= 0) {
        bucket = false;
      } else {
        chunk_src = std::max<uint64_t>(1, static_cast<uint64_t>(
            rec_words * waves / (1.25 * per_src * cold_share)));
        chunk_src = std::min<uint64_t>(chunk_src, n_src);
        // equal chunks
        const uint64_t n_chunks = (n_src + chunk_src - 1) / chunk_src;
        chunk_src = (n_src + n_chunks - 1) / n_chunks;
        uint64_t acc = 0;
        P.rbase[0] = 0;
        for (uint32_t k = 0; k < nb; ++k) {
          const double e = margin * static_cast<double>(chunk_src) * per_src * mass[k] / waves;
          uint64_t cap = static_cast<uint64_t>(e) + 2 * kFlush;
          cap = (cap + kFlush - 1) / kFlush * kFlush;
          acc += cap;
          if (acc > 0xFFFFFFFFull) return set_error(ctx, TGTP_E_UNSUPPORTED, "walk: record region");
          P.rbase[k + 1] = static_cast<uint32_t>(acc);
        }
        const uint64_t bytes = acc * waves * sizeof(uint32_t);
        rec = static_cast<uint32_t*>(ctx_buffer(ctx, kBufWalkRec, bytes));
        fill = static_cast<uint32_t*>(ctx_buffer(ctx, kBufWalkFill, waves * nb * sizeof(uint32_t)));
        if (!rec || !fill) bucket = false;
      }
      if (bucket) {
        P.rec = rec;
        P.fill = fill;
        Q.n = n;
        Q.hot = P.hot;
        Q.n_buckets = nb;
        Q.bshift = P.bshift;
        Q.passes = passes;
        Q.waves = static_cast<uint32_t>(waves);
        Q.rtot = P.rbase[nb];
        for (uint32_t k = 0; k <= nb; ++k) Q.rbase[k] = P.rbase[k];
        // about two workgroups per CU (one fits at a time: 128 KiB of LDS)
        Q.slices = static_cast<uint32_t>(std::max<uint64_t>(
            1, std::min<uint64_t>(waves, (2ull * ctx->num_cus + nb * passes - 1) / (nb * passes))));
      } else {
        // fall back to direct atomics with the plain kernel
        P.n_buckets = 0;
        chunk_src = n_src;
        fn = walk_kernel_ptr(packed, false);
      }
    }
    const size_t lds_launch = static_cast<size_t>(P.hot) * sizeof(uint32_t) +
                              (bucket ? sizeof(StageLds) : 0);
    if (!bucket) {
      TGTP_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWalkBlock,
                                                                     lds_launch));
      per_cu = std::max(1, std::min(per_cu, env_int("TGTP_WALK_BLOCKS_PER_CU", 64)));
    }
    const uint64_t grid_max = static_cast<uint64_t>(ctx->num_cus) * per_cu;
    // chunk c takes every n_chunks-th source from src_begin + c, so each
    // chunk's visits spread over the buckets like the whole run's
    const uint64_t n_chunks = (n_src + chunk_src - 1) / chunk_src;
    P.src_stride = static_cast<uint32_t>(n_chunks);
    for (uint64_t c = 0; c < n_chunks; ++c) {
      const uint64_t c_n = (n_src - c + n_chunks - 1) / n_chunks;
      P.src_begin = static_cast<uint32_t>(src_begin + c);
      P.total_bq = c_n * bps;
      const uint64_t tasks = (P.total_bq + 7) / 8;
      const uint64_t blocks_needed = (tasks + kWalkWaves - 1) / kWalkWaves;
      // bucketed: always the full resident grid, whose waves own the regions
      const uint64_t blocks = bucket ? grid_max
                                     : std::max<uint64_t>(1, std::min(blocks_needed, grid_max));
      P.stride_bq = blocks * kWalkWaves * 8ull;
      P.stride_src = static_cast<uint32_t>(P.stride_bq / bps);
      P.stride_rem = static_cast<uint32_t>(P.stride_bq % bps);
      const uint64_t* wadj = packed ? dg->walk_adj : nullptr;
      void* args[] = {const_cast<uint2**>(&dg->rowd), const_cast<uint32_t**>(&dg->adj),
                      const_cast<uint64_t**>(&wadj), &d_counters, &P};
      TGTP_HIP_CHECK(ctx, hipLaunchKernel(fn, dim3(static_cast<uint32_t>(blocks)),
                                          dim3(kWalkBlock), args, lds_launch, s));
      if (bucket) {
        bucket_count_kernel<<<nb * passes * Q.slices, kCountBlock, 0, s>>>(rec, fill, d_counters,
                                                                           Q);
        TGTP_HIP_CHECK(ctx, hipGetLastError());
      }
    }
  }
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[2], s));
  if ((rc = launch_walk_counter_sums(ctx, d_counters, n, d_sums))) return rc;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[3], s));
  TGTP_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev[3]));
  if (walk_ms) TGTP_HIP_CHECK(ctx, hipEventElapsedTime(walk_ms, ctx->ev[1], ctx->ev[2]));
  return TGTP_OK;
}
