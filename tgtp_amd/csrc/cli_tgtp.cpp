// tgtp drop-in (src/main.cpp:8-50): same arguments and validation, reads
// <folder>/<graph>/graph.txt and sorted_eigens_<omega>.txt, sets
// delta = 1/n, writes headerless float32 scores in CSR order (scores.bin).
// The scores come from tgtp_score on the GPU; TGTP_MODE selects the
// arithmetic (faithful by default, bit-exact with the reference);
// TGTP_CACHE=1 reads both text inputs through binary caches beside them.
//
// usage: tgtp <folder> <graph> <epsilon> <omega> <gamma> <outfile>
#include <climits>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <vector>

#include "cli_common.hpp"

int main(int argc, char** argv) {
  if (argc != 7) {
    std::cerr << "Usage: " << argv[0] << " <folder> <graph> <epsilon> <omega> <gamma> <outfile>\n";
    return 1;
  }
  try {
    auto number = [](const char* arg) {
      std::size_t used;
      const double v = std::stod(arg, &used);
      if (arg[used] != '\0' || !std::isfinite(v)) throw std::invalid_argument("invalid numeric argument");
      return v;
    };
    const double epsilon = number(argv[3]);
    const double omega = number(argv[4]), gamma = number(argv[5]);
    if (!(epsilon > 0 && epsilon < 1) || omega < 2 || omega > INT_MAX || omega != std::floor(omega) ||
        gamma < 1 || gamma > INT_MAX || gamma != std::floor(gamma))
      throw std::invalid_argument("require 0 < epsilon < 1, integer omega >= 2, integer gamma >= 1");
    const int32_t om = static_cast<int32_t>(omega);
    const auto dir = std::filesystem::path(argv[1]) / argv[2];
    cli::Csr csr;
    cli::load_graph((dir / "graph.txt").string(), csr);
    char err[512] = {0};
    if (tgtp_check_aesc_graph(&csr.g, om, err, sizeof(err))) throw std::runtime_error(err);
    std::vector<double> lam(om);
    std::vector<float> X(csr.g.n * static_cast<uint64_t>(om));
    const std::string epath = (dir / ("sorted_eigens_" + std::to_string(om) + ".txt")).string();
    cli::load_eigens(epath, csr.g.n, om, lam.data(), X.data());
    std::ofstream out(argv[6], std::ios::binary);
    if (!out) throw std::runtime_error("cannot open output file");
    cli::Ctx ctx(cli::device_from_env());
    const tgtp_graph gv = csr.view();
    const tgtp_eigens ev{om, lam.data(), X.data()};
    const tgtp_params pv{epsilon, 1.0 / static_cast<double>(csr.g.n), static_cast<int32_t>(gamma),
                         cli::mode_from_env()};
    std::vector<float> scores(csr.g.m);
    ctx.check(tgtp_score(ctx.c, &gv, &ev, &pv, scores.data(), nullptr), "tgtp_score");
    out.write(reinterpret_cast<const char*>(scores.data()),
              static_cast<std::streamsize>(scores.size() * sizeof(float)));
    out.close();
    if (!out) throw std::runtime_error("failed to write output file");
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "tgtp: " << e.what() << '\n';
    return 1;
  }
}
