// Compares tgtp_select::nth_element (the device selection) with the
// toolchain's std::nth_element on tie-heavy random inputs: the resulting
// permutations must be identical.  Exit status 0 = all equal.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "select.hpp"

int main(int argc, char** argv) {
  const int trials = argc > 1 ? std::atoi(argv[1]) : 20000;
  std::mt19937_64 rng(12345);
  long checked = 0;
  for (int t = 0; t < trials; ++t) {
    const int n = 1 + static_cast<int>(rng() % (t % 7 == 0 ? 5000 : 200));
    const int distinct = 1 + static_cast<int>(rng() % (t % 3 == 0 ? 4 : 64));
    std::vector<float> key(n + 17);
    std::vector<uint32_t> a(n), b;
    if (t % 4 != 3) {  // random keys with many ties
      for (auto& k : key) k = static_cast<float>(rng() % distinct) * 0.125f;
      for (int i = 0; i < n; ++i) a[i] = static_cast<uint32_t>(rng() % key.size());
    } else {  // organ-pipe / sawtooth / sorted keys and a median-of-three killer
              // order: deep partitioning, reaching the heap-selection branch
      for (int i = 0; i < static_cast<int>(key.size()); ++i)
        key[i] = (t % 12 == 3) ? static_cast<float>(i < n / 2 ? i : n - i)
                 : (t % 12 == 7) ? static_cast<float>(i % 17) : static_cast<float>(i);
      for (int i = 0; i < n; ++i) a[i] = static_cast<uint32_t>(i);
      for (int i = 0; i + 1 < n; i += 2) std::swap(a[i], a[n - 1 - i]);
    }
    b = a;
    std::vector<size_t> c(a.begin(), a.end());
    const int gamma = 1 + static_cast<int>(rng() % 40);
    const int rg = std::min(gamma, n);
    const float* kp = key.data();
    std::nth_element(c.begin(), c.begin() + (rg - 1), c.end(),
                     [kp](size_t x, size_t y) { return kp[x] > kp[y]; });
    tgtp_select::nth_element<uint32_t>(b.data(), rg - 1, n, tgtp_select::Greater<uint32_t>{kp});
    for (int i = 0; i < n; ++i) {
      if (b[i] != c[i]) {
        std::printf("mismatch trial %d n=%d gamma=%d at %d\n", t, n, gamma, i);
        return 1;
      }
    }
    ++checked;
  }
  std::printf("ok %ld\n", checked);
  return 0;
}
