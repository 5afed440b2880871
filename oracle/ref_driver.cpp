// ref_driver.cpp -- a caller of the reference's own code, compiled against
// the reference sources where they lie (see oracle/Makefile, target `ref`).
// It is the strongest available check of the oracle: aesc_baseline.h holds
// the same tau / diffusion / top-gamma / chi / walk-count / symmetrization
// code as aesc.h (aesc_baseline.h:99-362 vs aesc.h:129-551) with a scalar
// MT19937 walk, and it compiles with plain g++ (aesc.h needs Intel SVML's
// _mm256_rem_epu32, which this image lacks, so aesc.h itself is not built).
//
// usage: ref_tgtp_mt <graph.txt> <eigens.txt> <omega> <epsilon> <gamma> <out_prefix>
// writes <out_prefix>.idx.u32 .adj.u32 (graph.h load_txt), .lam.f64 .vec.f32
// (EigenVecs::load_txt), .taus.i32 (calculate_taus) and .scores.f32 (tgtp()).
#include "aesc_baseline.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <string>

static void dump(const std::string& path, const void* p, size_t bytes) {
  std::ofstream out(path, std::ios::binary);
  out.write(static_cast<const char*>(p), static_cast<std::streamsize>(bytes));
  if (!out) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(1);
  }
}

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s graph eigens omega eps gamma out_prefix\n", argv[0]);
    return 1;
  }
  try {
    const int omega = std::atoi(argv[3]);
    AESCGraph graph(argv[1], argv[2], omega);
    AESCConfig config{};
    config.epsilon = std::atof(argv[4]);
    config.omega = omega;
    config.gamma = std::atoi(argv[5]);
    config.delta = 1.0 / graph.n;  // as main.cpp:36
    const std::string pre = argv[6];
    dump(pre + ".idx.u32", graph.idx, (graph.n + 1) * sizeof(uint32_t));
    dump(pre + ".adj.u32", graph.adj, graph.m * sizeof(uint32_t));
    dump(pre + ".lam.f64", graph.eigens.values, omega * sizeof(double));
    dump(pre + ".vec.f32", graph.eigens.vectors, graph.n * omega * sizeof(float));
    AESC algo(graph, config);
    std::unique_ptr<int[]> taus(algo.calculate_taus(config.epsilon / 2.0));
    dump(pre + ".taus.i32", taus.get(), graph.m * sizeof(int));
    std::unique_ptr<float[]> scores(algo.tgtp());
    dump(pre + ".scores.f32", scores.get(), graph.m * sizeof(float));
    graph.free();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ref_tgtp_mt: %s\n", e.what());
    return 1;
  }
  return 0;
}
