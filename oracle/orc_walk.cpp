// orc_walk.cpp -- CPU restatement of the SABA walks (test infrastructure
// only; see oracle.h).  Lanes of an 8-wide bouquet are plain arrays here:
// the restatement follows the reference's per-lane arithmetic, not its AVX2
// data movement.
#include "oracle.h"

#include <omp.h>

#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

namespace {
constexpr int kLanes = 8;  // BLOCKSIZE, aesc.h:247; rw_experiment_saba.cpp:70

inline uint32_t deg(const uint32_t* idx, uint32_t v) { return idx[v + 1] - idx[v]; }

// high 32 bits of a 32x32 product: the multiply-high step of
// rw_experiment_saba.cpp:93-96 / aesc.h:295-302.
inline uint32_t mulhi32(uint32_t a, uint32_t b) {
  return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) >> 32);
}
}  // namespace

extern "C" uint32_t orc_fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

extern "C" uint32_t orc_splitmix32(uint32_t* state) {
  *state += 0x9e3779b9u;
  uint32_t z = *state ^ (*state >> 16);
  z *= 0x21f0aaadu;
  z ^= z >> 15;
  z *= 0x735a2d97u;
  return z ^ (z >> 15);
}

// rw_experiment_saba.cpp:65-145.
extern "C" void orc_saba_walk_source(const uint32_t* idx, const uint32_t* adj,
                                     uint32_t src, uint32_t depth,
                                     uint64_t n_walks, uint32_t* counters) {
  if (n_walks == 0) return;
  // UINT32_MAX / uint32_t(n_walk)   (:71)
  const uint32_t step = 0xFFFFFFFFu / static_cast<uint32_t>(n_walks);
  for (uint64_t ii = 0; ii < n_walks; ii += kLanes) {
    uint32_t cur[kLanes], sec[kLanes], st[kLanes];
    for (int b = 0; b < kLanes; ++b) {
      cur[b] = src;
      sec[b] = (static_cast<uint32_t>(ii) + static_cast<uint32_t>(b + 1)) * step;  // :75-78
      st[b] = sec[b];                                                                // :79
    }
    // A walker at a vertex without neighbours (reachable only in a
    // non-symmetric CSR, where the reference would divide by zero) stops:
    // it no longer moves, counts, or joins its bouquet's zero check.  In a
    // graph storing both directions every walker stays live.
    bool live[kLanes];
    uint32_t hop = 0;
    for (; hop < depth; ++hop) {
      bool any_zero = false;  // :83-85, checked before every hashed hop
      for (int b = 0; b < kLanes; ++b) {
        live[b] = (hop == 0 || live[b]) && deg(idx, cur[b]) != 0;
        any_zero |= live[b] && sec[b] == 0;
      }
      if (any_zero) break;
      for (int b = 0; b < kLanes; ++b) {
        if (!live[b]) continue;
        const uint32_t d = deg(idx, cur[b]);
        const uint32_t k = mulhi32(sec[b] ^ orc_fmix32(cur[b]), d);
        cur[b] = adj[idx[cur[b]] + k];
        sec[b] *= d;  // :106
        counters[cur[b]] += 1u;
      }
    }
    // Whole-bouquet SplitMix32 fallback for the remaining hops (:117-143).
    for (; hop < depth; ++hop) {
      for (int b = 0; b < kLanes; ++b) {
        live[b] = live[b] && deg(idx, cur[b]) != 0;
        if (!live[b]) continue;
        const uint32_t r = orc_splitmix32(&st[b]);
        const uint32_t k = r % deg(idx, cur[b]);
        cur[b] = adj[idx[cur[b]] + k];
        counters[cur[b]] += 1u;
      }
    }
  }
}

// Driver loop of rw_experiment_saba.cpp:190-207 with uint32 counters.
extern "C" int orc_saba_walks(uint64_t n, const uint32_t* idx,
                              const uint32_t* adj, uint32_t src_begin,
                              uint32_t src_end, uint32_t depth,
                              uint64_t n_walks, uint32_t* counters, int threads,
                              int chunk) {
  if (n_walks % kLanes != 0 || n_walks > 0xFFFFFFFFull) return -1;
  if (src_end > n) src_end = static_cast<uint32_t>(n);
  if (chunk <= 0) chunk = 1024;
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
  {
    std::vector<uint32_t> local(n, 0u);
#pragma omp for schedule(dynamic, chunk)
    for (int64_t s = src_begin; s < static_cast<int64_t>(src_end); ++s) {
      const uint32_t src = static_cast<uint32_t>(s);
      if (deg(idx, src) == 0) continue;
      orc_saba_walk_source(idx, adj, src, depth, n_walks, local.data());
    }
#pragma omp critical
    for (uint64_t v = 0; v < n; ++v) counters[v] += local[v];
  }
  return 0;
}

// aesc.h:241-391.  Both walkers' grid secrets start equal; the hashed phase
// ends for the bouquet pair when any lane of either secret is zero; the
// fallback draws both walkers' numbers from the one per-lane state.
extern "C" float orc_walk2_saba(const uint32_t* idx, const uint32_t* adj,
                                uint32_t s, uint32_t v, uint64_t nr,
                                uint32_t len, const float* pvec) {
  const uint32_t step = static_cast<uint32_t>(0xFFFFFFFFull / nr);  // :248
  float x[kLanes] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint64_t i = 0; i < nr; i += kLanes) {
    uint32_t c1[kLanes], c2[kLanes], s1[kLanes], s2[kLanes], st[kLanes];
    for (int b = 0; b < kLanes; ++b) {
      c1[b] = s;
      c2[b] = v;
      s1[b] = (static_cast<uint32_t>(i) + static_cast<uint32_t>(b + 1)) * step;
      s2[b] = s1[b];
      st[b] = s1[b];
    }
    uint32_t hop = 0;
    for (; hop < len; ++hop) {
      bool any_zero = false;
      for (int b = 0; b < kLanes; ++b) any_zero |= (s1[b] == 0) | (s2[b] == 0);
      if (any_zero) break;
      for (int b = 0; b < kLanes; ++b) {
        const uint32_t d1 = deg(idx, c1[b]);
        const uint32_t d2 = deg(idx, c2[b]);
        const uint32_t k1 = mulhi32(s1[b] ^ orc_fmix32(c1[b]), d1);
        const uint32_t k2 = mulhi32(s2[b] ^ orc_fmix32(c2[b]), d2);
        c1[b] = adj[idx[c1[b]] + k1];
        c2[b] = adj[idx[c2[b]] + k2];
        s1[b] *= d1;
        s2[b] *= d2;
        x[b] += pvec[c1[b]] - pvec[c2[b]];
      }
    }
    for (; hop < len; ++hop) {
      for (int b = 0; b < kLanes; ++b) {
        const uint32_t r1 = orc_splitmix32(&st[b]);
        const uint32_t r2 = orc_splitmix32(&st[b]);
        const uint32_t k1 = r1 % deg(idx, c1[b]);
        const uint32_t k2 = r2 % deg(idx, c2[b]);
        c1[b] = adj[idx[c1[b]] + k1];
        c2[b] = adj[idx[c2[b]] + k2];
        x[b] += pvec[c1[b]] - pvec[c2[b]];
      }
    }
  }
  float sum = 0;  // :386-390, lanes 0..7 in order
  for (int b = 0; b < kLanes; ++b) sum += x[b];
  return sum;
}

// aesc_baseline.h:363-379: scalar 2-way walk drawing from one generator.
float orc_walk2_mt(const uint32_t* idx, const uint32_t* adj, uint32_t s,
                   uint32_t v, uint64_t nr, uint64_t len, const float* pvec,
                   std::mt19937& gen) {
  float x = 0;
  for (uint64_t i = 0; i < nr; ++i) {
    uint32_t c1 = s, c2 = v;
    for (uint32_t hop = 0; hop < len; ++hop) {
      const uint32_t d1 = deg(idx, c1);
      const uint32_t k1 = gen() % d1;
      c1 = adj[idx[c1] + k1];
      const uint32_t d2 = deg(idx, c2);
      const uint32_t k2 = gen() % d2;
      c2 = adj[idx[c2] + k2];
      x += (pvec[c1] - pvec[c2]);
    }
  }
  return x;
}

// ---- MT19937 comparator walks -------------------------------------------
// std::mt19937 itself (the reference's generator); the C++ standard fixes it
// ([rand.predef]: the 10000th output of a default-seeded mt19937 is
// 4123659995), so orc_mt19937_nth checks the library.
extern "C" uint32_t orc_mt19937_nth(uint32_t seed, uint64_t k) {
  std::mt19937 rng(seed);
  uint32_t x = 0;
  for (uint64_t i = 0; i < k; ++i) x = rng();
  return x;
}

// rw_experiment.cpp:34-47 (variant 0: walk i takes draws i*depth .. +depth-1)
// and rw_experiment_simd.cpp:44-82 (variant 1: bouquets of 8, each hop draws
// the 8 lanes' numbers in lane order), with the per-source reseeding of
// rw_experiment.cpp:99-104 and uint32 counters.
extern "C" int orc_mt_walks(uint64_t n, const uint32_t* idx, const uint32_t* adj,
                            uint32_t src_begin, uint32_t src_end, uint32_t depth,
                            uint64_t n_walks, int variant, uint32_t* counters, int threads) {
  if (variant != 0 && variant != 1) return -1;
  if (variant == 1 && n_walks % 8 != 0) return -1;
  if (src_end > n) src_end = static_cast<uint32_t>(n);
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
  {
    std::vector<uint32_t> local(n, 0u);
    std::mt19937 rng(42u);
#pragma omp for schedule(dynamic, 1024)
    for (int64_t s = src_begin; s < static_cast<int64_t>(src_end); ++s) {
      const uint32_t src = static_cast<uint32_t>(s);
      if (deg(idx, src) == 0) continue;
      rng.seed(42u + src);
      // A walker at a vertex without neighbours (reachable only in a
      // non-symmetric CSR; the reference divides by zero there) stays and
      // stops counting; its draws are still taken, so every other walk
      // keeps its place in the stream.
      if (variant == 0) {
        for (uint64_t i = 0; i < n_walks; ++i) {
          uint32_t cur = src;
          for (uint32_t h = 0; h < depth; ++h) {
            const uint32_t r = rng();
            const uint32_t d = deg(idx, cur);
            if (d == 0) continue;
            cur = adj[idx[cur] + r % d];
            ++local[cur];
          }
        }
      } else {
        for (uint64_t ii = 0; ii < n_walks; ii += 8) {
          uint32_t cur[8];
          for (int j = 0; j < 8; ++j) cur[j] = src;
          for (uint32_t h = 0; h < depth; ++h) {
            uint32_t rnd[8];
            for (int j = 0; j < 8; ++j) rnd[j] = rng();
            for (int j = 0; j < 8; ++j) {
              const uint32_t d = deg(idx, cur[j]);
              if (d == 0) continue;
              cur[j] = adj[idx[cur[j]] + rnd[j] % d];
              ++local[cur[j]];
            }
          }
        }
      }
    }
#pragma omp critical
    for (uint64_t v = 0; v < n; ++v) counters[v] += local[v];
  }
  return 0;
}
