// orc_walk_avx2.cpp -- the SABA 1-way walks on the host with AVX2: the CPU
// baseline the GPU walk engine is measured against (test infrastructure;
// see oracle.h).  Same results as orc_saba_walks (rw_experiment_saba.cpp:
// 65-145 with uint32 counters), bit for bit: the 8 walkers of a bouquet are
// the 8 lanes of one 256-bit register, row bounds and neighbours are
// gathered, the multiply-high step is two 32x32->64 multiplies (even and odd
// lanes), and counters are bumped lane by lane in lane order.  A bouquet
// leaves the vector loop at its first zero secret (the whole-bouquet
// SplitMix32 fallback, :116-143, which the reference computes with a lane-wise
// remainder) or at a vertex without neighbours (non-symmetric CSR), and the
// scalar rule of orc_saba_walk_source finishes it from the same state.
#include "oracle.h"

#include <immintrin.h>
#include <omp.h>

#include <cstdint>
#include <vector>

namespace {

constexpr int kLanes = 8;

inline uint32_t splitmix32(uint32_t* state) { return orc_splitmix32(state); }

// Scalar continuation from hop `hop` with every walker still live: the rules
// of orc_saba_walk_source (hashed hops until a zero secret, then the
// SplitMix32 fallback from the grid secrets; a walker at a vertex without
// neighbours stops).
void finish_bouquet(const uint32_t* idx, const uint32_t* adj, uint32_t depth, uint32_t hop,
                    uint32_t* cur, uint32_t* sec, const uint32_t* sec0, uint32_t* counters) {
  bool live[kLanes];
  for (int b = 0; b < kLanes; ++b) live[b] = true;
  uint32_t st[kLanes];
  for (int b = 0; b < kLanes; ++b) st[b] = sec0[b];
  for (; hop < depth; ++hop) {
    bool any_zero = false;
    for (int b = 0; b < kLanes; ++b) {
      live[b] = live[b] && idx[cur[b] + 1] != idx[cur[b]];
      any_zero |= live[b] && sec[b] == 0;
    }
    if (any_zero) break;
    for (int b = 0; b < kLanes; ++b) {
      if (!live[b]) continue;
      const uint32_t d = idx[cur[b] + 1] - idx[cur[b]];
      const uint32_t k = static_cast<uint32_t>(
          (static_cast<uint64_t>(sec[b] ^ orc_fmix32(cur[b])) * d) >> 32);
      cur[b] = adj[idx[cur[b]] + k];
      sec[b] *= d;
      counters[cur[b]] += 1u;
    }
  }
  for (; hop < depth; ++hop) {
    for (int b = 0; b < kLanes; ++b) {
      live[b] = live[b] && idx[cur[b] + 1] != idx[cur[b]];
      if (!live[b]) continue;
      const uint32_t r = splitmix32(&st[b]);
      cur[b] = adj[idx[cur[b]] + r % (idx[cur[b] + 1] - idx[cur[b]])];
      counters[cur[b]] += 1u;
    }
  }
}

__attribute__((target("avx2"))) inline __m256i fmix32_v(__m256i h) {
  h = _mm256_xor_si256(h, _mm256_srli_epi32(h, 16));
  h = _mm256_mullo_epi32(h, _mm256_set1_epi32(static_cast<int>(0x85ebca6bu)));
  h = _mm256_xor_si256(h, _mm256_srli_epi32(h, 13));
  h = _mm256_mullo_epi32(h, _mm256_set1_epi32(static_cast<int>(0xc2b2ae35u)));
  return _mm256_xor_si256(h, _mm256_srli_epi32(h, 16));
}

// high 32 bits of each lane's 32x32 product
__attribute__((target("avx2"))) inline __m256i mulhi_v(__m256i a, __m256i b) {
  const __m256i even = _mm256_srli_epi64(_mm256_mul_epu32(a, b), 32);
  const __m256i odd = _mm256_mul_epu32(_mm256_srli_epi64(a, 32), _mm256_srli_epi64(b, 32));
  return _mm256_blend_epi32(even, odd, 0xAA);
}

__attribute__((target("avx2"))) void walk_source_avx2(const uint32_t* idx, const uint32_t* adj,
                                                      uint32_t src, uint32_t depth,
                                                      uint32_t n_walks, uint32_t* counters) {
  const uint32_t step = 0xFFFFFFFFu / n_walks;
  const int* idx_i = reinterpret_cast<const int*>(idx);
  const int* adj_i = reinterpret_cast<const int*>(adj);
  const __m256i lane1 = _mm256_setr_epi32(1, 2, 3, 4, 5, 6, 7, 8);
  const __m256i one = _mm256_set1_epi32(1);
  const __m256i zero = _mm256_setzero_si256();
  const __m256i vstep = _mm256_set1_epi32(static_cast<int>(step));
  for (uint32_t ii = 0; ii < n_walks; ii += kLanes) {
    // secret of lane b: (ii + b + 1) * step (rw_experiment_saba.cpp:75-78)
    const __m256i sec0 = _mm256_mullo_epi32(
        _mm256_add_epi32(_mm256_set1_epi32(static_cast<int>(ii)), lane1), vstep);
    __m256i sec = sec0;
    __m256i cur = _mm256_set1_epi32(static_cast<int>(src));
    uint32_t hop = 0;
    for (; hop < depth; ++hop) {
      const __m256i lo = _mm256_i32gather_epi32(idx_i, cur, 4);
      const __m256i hi = _mm256_i32gather_epi32(idx_i, _mm256_add_epi32(cur, one), 4);
      const __m256i d = _mm256_sub_epi32(hi, lo);
      const int stop = _mm256_movemask_ps(_mm256_castsi256_ps(_mm256_or_si256(
          _mm256_cmpeq_epi32(sec, zero), _mm256_cmpeq_epi32(d, zero))));
      if (stop) break;
      const __m256i k = mulhi_v(_mm256_xor_si256(sec, fmix32_v(cur)), d);
      cur = _mm256_i32gather_epi32(adj_i, _mm256_add_epi32(lo, k), 4);
      sec = _mm256_mullo_epi32(sec, d);
      alignas(32) uint32_t c[kLanes];
      _mm256_store_si256(reinterpret_cast<__m256i*>(c), cur);
      for (int b = 0; b < kLanes; ++b) counters[c[b]] += 1u;
    }
    if (hop < depth) {
      alignas(32) uint32_t c[kLanes], s[kLanes], s0[kLanes];
      _mm256_store_si256(reinterpret_cast<__m256i*>(c), cur);
      _mm256_store_si256(reinterpret_cast<__m256i*>(s), sec);
      _mm256_store_si256(reinterpret_cast<__m256i*>(s0), sec0);
      finish_bouquet(idx, adj, depth, hop, c, s, s0, counters);
    }
  }
}

}  // namespace

extern "C" int orc_saba_walks_avx2(uint64_t n, const uint32_t* idx, const uint32_t* adj,
                                   uint32_t src_begin, uint32_t src_end, uint32_t depth,
                                   uint64_t n_walks, uint32_t* counters, int threads,
                                   int chunk) {
  if (!__builtin_cpu_supports("avx2")) return -2;
  // gathers use signed 32-bit indices
  if (n_walks % kLanes != 0 || n_walks > 0xFFFFFFFFull || n >= (1ull << 31) ||
      idx[n] >= (1u << 31))
    return -1;
  if (src_end > n) src_end = static_cast<uint32_t>(n);
  if (chunk <= 0) chunk = 1024;
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
  {
    std::vector<uint32_t> local(n, 0u);
#pragma omp for schedule(dynamic, chunk)
    for (int64_t s = src_begin; s < static_cast<int64_t>(src_end); ++s) {
      const uint32_t src = static_cast<uint32_t>(s);
      if (idx[src + 1] == idx[src] || n_walks == 0) continue;
      walk_source_avx2(idx, adj, src, depth, static_cast<uint32_t>(n_walks), local.data());
    }
#pragma omp critical
    for (uint64_t v = 0; v < n; ++v) counters[v] += local[v];
  }
  return 0;
}
