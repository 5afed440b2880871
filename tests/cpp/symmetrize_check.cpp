// tgtp_host::symmetrize_sources (tgtp_amd/csrc/host_tgtp.cpp) on seeded
// random CSRs: for sorted rows holding both directions of every edge
// (duplicates and self-loops included) the map equals a direct restatement
// of aesc.h:538-545 -- rev_preds[first position of u in row v] takes
// preds[pos(u, v)], rows u ascending, positions ascending, the last write
// winning -- and the return code is 0; an unsorted row gives 1, an edge
// without its reverse 2 -- on small graphs and on graphs large enough for the
// host threads to split the rows.
// usage: symmetrize_check <cases>; prints "ok <cases>" on success.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "host_tgtp.hpp"

namespace {

struct Csr {
  std::vector<uint32_t> idx, adj;
};

Csr from_pairs(uint32_t n, const std::vector<std::pair<uint32_t, uint32_t>>& pairs) {
  std::vector<std::vector<uint32_t>> rows(n);
  for (auto [a, b] : pairs) {
    rows[a].push_back(b);
    rows[b].push_back(a);
  }
  Csr g;
  g.idx.push_back(0);
  for (auto& r : rows) {
    std::sort(r.begin(), r.end());
    g.adj.insert(g.adj.end(), r.begin(), r.end());
    g.idx.push_back(static_cast<uint32_t>(g.adj.size()));
  }
  return g;
}

std::vector<int32_t> direct(const Csr& g) {
  const uint32_t n = static_cast<uint32_t>(g.idx.size() - 1);
  std::vector<int32_t> src(g.adj.size(), -1);
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t p = g.idx[u]; p < g.idx[u + 1]; ++p) {
      const uint32_t v = g.adj[p];
      uint32_t q = g.idx[v];
      while (g.adj[q] != u) ++q;  // first u in row v (present: both directions stored)
      src[q] = static_cast<int32_t>(p);
    }
  return src;
}

}  // namespace

int main(int argc, char** argv) {
  const int cases = argc > 1 ? std::atoi(argv[1]) : 200;
  std::mt19937_64 rng(77);
  for (int c = 0; c < cases; ++c) {
    const uint32_t n = 2 + static_cast<uint32_t>(rng() % 40);
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    const int e = 1 + static_cast<int>(rng() % (3 * n));
    for (int k = 0; k < e; ++k) {
      const uint32_t a = static_cast<uint32_t>(rng() % n), b = static_cast<uint32_t>(rng() % n);
      pairs.emplace_back(a, b);
      if (rng() % 5 == 0) pairs.emplace_back(a, b);  // duplicate
    }
    Csr g = from_pairs(n, pairs);
    std::vector<int32_t> src;
    int rc = tgtp_host::symmetrize_sources(g.idx.data(), g.adj.data(), n, g.adj.size(), src);
    if (rc != 0 || src != direct(g)) {
      std::printf("case %d: symmetric graph rc %d or map differs\n", c, rc);
      return 1;
    }
    // an edge without its reverse: redirect one entry of a row of length 1
    // or more to a vertex whose row lacks that row's vertex
    Csr h = g;
    bool made = false;
    for (uint32_t u = 0; u < n && !made; ++u)
      for (uint32_t p = h.idx[u]; p < h.idx[u + 1] && !made; ++p)
        for (uint32_t w = 0; w < n && !made; ++w) {
          const bool w_has_u = std::count(h.adj.begin() + h.idx[w], h.adj.begin() + h.idx[w + 1], u) > 0;
          const bool sorted_ok = (p == h.idx[u] || h.adj[p - 1] <= w) &&
                                 (p + 1 == h.idx[u + 1] || w <= h.adj[p + 1]);
          if (!w_has_u && w != u && sorted_ok) {
            h.adj[p] = w;
            made = true;
          }
        }
    if (made) {
      rc = tgtp_host::symmetrize_sources(h.idx.data(), h.adj.data(), n, h.adj.size(), src);
      if (rc != 2) {
        std::printf("case %d: missing reverse gave %d\n", c, rc);
        return 1;
      }
    }
    // an unsorted row
    for (uint32_t u = 0; u < n; ++u)
      if (g.idx[u + 1] - g.idx[u] >= 2 && g.adj[g.idx[u]] != g.adj[g.idx[u + 1] - 1]) {
        Csr k = g;
        std::swap(k.adj[k.idx[u]], k.adj[k.idx[u + 1] - 1]);
        rc = tgtp_host::symmetrize_sources(k.idx.data(), k.adj.data(), n, k.adj.size(), src);
        if (rc != 1) {
          std::printf("case %d: unsorted row gave %d\n", c, rc);
          return 1;
        }
        break;
      }
  }
  // graphs large enough (over 2^20 CSR entries) that the host threads split
  // the rows
  for (int c = 0; c < 2; ++c) {
    const uint32_t n = 100000 + static_cast<uint32_t>(rng() % 20000);
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (uint32_t k = 0; k < 6 * n; ++k) {
      const uint32_t a = static_cast<uint32_t>(rng() % n);
      const uint32_t b = rng() % 64 == 0 ? static_cast<uint32_t>(rng() % 64)  // hubs
                                         : static_cast<uint32_t>(rng() % n);
      pairs.emplace_back(a, b);
      if (rng() % 7 == 0) pairs.emplace_back(b, a);  // duplicate
    }
    const Csr g = from_pairs(n, pairs);
    if (g.adj.size() < (1u << 20)) {
      std::printf("large case %d: only %zu entries\n", c, g.adj.size());
      return 1;
    }
    std::vector<int32_t> src;
    const int rc = tgtp_host::symmetrize_sources(g.idx.data(), g.adj.data(), n, g.adj.size(), src);
    if (rc != 0 || src != direct(g)) {
      std::printf("large case %d: rc %d or map differs\n", c, rc);
      return 1;
    }
  }
  std::printf("ok %d\n", cases);
  return 0;
}
