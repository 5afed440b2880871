// aesc_example.cpp -- a caller written against the AESC interface of the
// reference (AESCGraph / AESCConfig / AESC::tgtp(), src/aesc.h:83-127), built
// on the drop-in header include/tgtp_aesc.hpp instead of aesc.h.  It takes the
// arguments of the tgtp binary (src/main.cpp:8-32) and writes the same
// scores.bin, so the two routes to the GPU can be compared.
//
// usage: aesc_example <folder> <graph> <epsilon> <omega> <gamma> <outfile>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>

#include "tgtp_aesc.hpp"

int main(int argc, char** argv) {
  if (argc != 7) {
    std::cerr << "usage: " << argv[0] << " <folder> <graph> <epsilon> <omega> <gamma> <outfile>\n";
    return 1;
  }
  try {
    AESCConfig config{};
    config.strFolder = argv[1];
    config.strGraph = argv[2];
    config.epsilon = std::stod(argv[3]);
    config.omega = std::stoi(argv[4]);
    config.gamma = std::stoi(argv[5]);
    const auto dir = std::filesystem::path(config.strFolder) / config.strGraph;
    const std::string eigen_file = "sorted_eigens_" + std::to_string(config.omega) + ".txt";
    AESCGraph graph((dir / "graph.txt").string(), (dir / eigen_file).string(),
                    static_cast<size_t>(config.omega));
    config.delta = 1.0 / static_cast<double>(graph.n);
    AESC algo(graph, config);
    std::unique_ptr<float[]> scores(algo.tgtp());
    std::ofstream out(argv[6], std::ios::binary);
    out.write(reinterpret_cast<const char*>(scores.get()),
              static_cast<std::streamsize>(graph.m * sizeof(float)));
    out.close();
    if (!out) throw std::runtime_error("failed to write output file");
    graph.free();
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "aesc_example: " << e.what() << '\n';
    return 1;
  }
}
