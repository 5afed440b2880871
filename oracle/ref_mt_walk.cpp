// ref_mt_walk.cpp -- a caller of the reference's own scalar_randomwalk
// (rw_experiment.cpp:34-47), compiled against the reference source where it
// lies (oracle/Makefile, target ref; its main() is renamed so that this
// driver can call the static walk function).  It pins the oracle's MT19937
// walks: per source, std::mt19937 reseeded with 42 + source as the
// reference's driver does (rw_experiment.cpp:99-104), counts accumulated in
// the reference's float counters (exact below 2^24 per vertex).
//
// usage: ref_mt_walk <graph.txt> <depth> <n_walks> <out.f32>
#define main rw_experiment_main
#include "rw_experiment.cpp"
#undef main

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s graph.txt depth n_walks out.f32\n", argv[0]);
    return 1;
  }
  try {
    fuser::graph_t<uint32_t, uint32_t> g;
    g.load_txt(argv[1]);
    const size_t depth = std::strtoull(argv[2], nullptr, 10);
    const size_t n_walks = std::strtoull(argv[3], nullptr, 10);
    std::vector<float> pvec(g.n, 0.0f);
    std::mt19937 rng(42u);
    for (size_t s = 0; s < g.n; ++s) {
      const uint32_t src = static_cast<uint32_t>(s);
      if (g.degree(src) == 0) continue;
      rng.seed(42u + src);
      scalar_randomwalk(src, depth, n_walks, pvec.data(), g, rng);
    }
    std::ofstream out(argv[4], std::ios::binary);
    out.write(reinterpret_cast<const char*>(pvec.data()),
              static_cast<std::streamsize>(pvec.size() * sizeof(float)));
    g.free();
    if (!out) return 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ref_mt_walk: %s\n", e.what());
    return 1;
  }
  return 0;
}
