// Walk experiment drop-ins, one source built three ways (TGTP_RW_VARIANT):
//   0 rw_experiment_saba  (src/rw_experiment_saba.cpp:147-247) tgtp_saba_walks
//   1 rw_experiment       (src/rw_experiment.cpp:49-152)      tgtp_mt_walks, regular
//   2 rw_experiment_simd  (src/rw_experiment_simd.cpp:84-191) tgtp_mt_walks, simd
// The same arguments and validation, the same 14-column TSV row on stdout
// and the same results/rw/<variant>_<dataset>_d<depth>_w<n_walks>.txt
// key/value file; the walks run on the GPU.
//
// usage: <binary> <datasets_folder> <dataset> <depth> <n_walks> [threads]
// t_load_s covers parsing and the upload to HBM; t_rw_s covers counter
// clearing, the walks and the counter reduction (the reference's window).
#include <chrono>
#include <cstdlib>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>

#include "cli_common.hpp"

namespace fs = std::filesystem;
#ifndef TGTP_RW_VARIANT
#define TGTP_RW_VARIANT 0
#endif
#if TGTP_RW_VARIANT == 0
static const char* const kVariant = "saba";
#elif TGTP_RW_VARIANT == 1
static const char* const kVariant = "regular";
#else
static const char* const kVariant = "simd";
#endif

int main(int argc, char** argv) {
  if (argc < 5) {
    std::cerr << "Usage: " << argv[0]
              << " <datasets_folder> <dataset> <depth> <n_walks> [threads]\n";
    return 1;
  }
  const std::string folder = argv[1];
  const std::string dataset = argv[2];
  const size_t depth = std::strtoull(argv[3], nullptr, 10);
  const size_t n_walks = std::strtoull(argv[4], nullptr, 10);
#if TGTP_RW_VARIANT == 0
  if (n_walks % 8 != 0 || n_walks > UINT32_MAX || depth > UINT32_MAX) {
    std::cerr << "n_walks must be a multiple of 8 within the 32-bit grid domain\n";
    return 1;
  }
#elif TGTP_RW_VARIANT == 2
  if (n_walks % 8 != 0) {  // rw_experiment_simd.cpp:97-100
    std::cerr << "n_walks must be a multiple of 8\n";
    return 1;
  }
#endif
  if (n_walks > UINT32_MAX || depth > UINT32_MAX) {
    std::cerr << "depth and n_walks must fit 32 bits\n";
    return 1;
  }
  int threads = (argc > 5) ? std::atoi(argv[5]) : 1;  // host threads; reported as given
  if (threads < 1) threads = 1;
  try {
    cli::Ctx ctx(cli::device_from_env());
    const fs::path graph_path = fs::path(folder) / dataset / "graph.txt";
    cli::Csr csr;
    const auto t0 = std::chrono::high_resolution_clock::now();
    cli::load_graph(graph_path.string(), csr);
    const tgtp_graph gv = csr.view();
    tgtp_dgraph* dg = nullptr;
    ctx.check(tgtp_graph_upload(ctx.c, &gv, &dg), "tgtp_graph_upload");
    const auto t1 = std::chrono::high_resolution_clock::now();
    const double t_load = std::chrono::duration<double>(t1 - t0).count();
    const size_t n = csr.g.n, m = csr.g.m;

    uint64_t vc = 0, cs = 0;
    const auto t2 = std::chrono::high_resolution_clock::now();
    if (n_walks > 0) {
#if TGTP_RW_VARIANT == 0
      ctx.check(tgtp_saba_walks(ctx.c, dg, 0, static_cast<uint32_t>(n), static_cast<uint32_t>(depth),
                                static_cast<uint32_t>(n_walks), nullptr, &vc, &cs, nullptr),
                "tgtp_saba_walks");
#else
      ctx.check(tgtp_mt_walks(ctx.c, dg, 0, static_cast<uint32_t>(n), static_cast<uint32_t>(depth),
                              static_cast<uint32_t>(n_walks),
                              TGTP_RW_VARIANT == 1 ? TGTP_WALK_MT_REGULAR : TGTP_WALK_MT_SIMD,
                              nullptr, &vc, &cs, nullptr),
                "tgtp_mt_walks");
#endif
    }
    const auto t3 = std::chrono::high_resolution_clock::now();
    tgtp_graph_free(ctx.c, dg);
    const double t_rw = std::chrono::duration<double>(t3 - t2).count();
    const double total = std::chrono::duration<double>(t3 - t0).count();
    const double total_walks = static_cast<double>(n) * static_cast<double>(n_walks) *
                               static_cast<double>(depth);
    const double walks_per_sec = t_rw > 0.0 ? total_walks / t_rw : 0.0;
    // the reference accumulates these in double (rw_experiment_saba.cpp:188,
    // rw_experiment.cpp:85)
    const double visit_count = static_cast<double>(vc);
    const double checksum = static_cast<double>(cs);

    fs::create_directories("results/rw");
    const std::string resfile = "results/rw/" + std::string(kVariant) + "_" + dataset + "_d" +
                                std::to_string(depth) + "_w" + std::to_string(n_walks) + ".txt";
    {
      std::ofstream out(resfile);
      out << "variant " << kVariant << "\n"
          << "dataset " << dataset << "\n"
          << "depth " << depth << "\n"
          << "n_walks " << n_walks << "\n"
          << "threads " << threads << "\n"
          << "n_vertices " << n << "\n"
          << "n_edges " << m << "\n"
          << "t_load_s " << t_load << "\n"
          << "t_rw_s " << t_rw << "\n"
          << "total_s " << total << "\n"
          << "total_walks " << total_walks << "\n"
          << "walks_per_sec " << walks_per_sec << "\n"
          << "visit_count " << visit_count << "\n"
          << "checksum " << checksum << "\n";
      // TGTP_RW_EXACT=1 appends the exact integer sums (the double lines
      // above round beyond 2^53); off by default, so the file has the
      // reference's keys (rw_experiment_saba.cpp:224-238)
      const char* ex = std::getenv("TGTP_RW_EXACT");
      if (ex && std::string(ex) == "1")
        out << "visit_count_exact " << vc << "\n"
            << "checksum_exact " << cs << "\n";
    }
    std::cout << kVariant << "\t" << dataset << "\t" << depth << "\t" << n_walks << "\t" << threads
              << "\t" << n << "\t" << m << "\t" << t_load << "\t" << t_rw << "\t" << total << "\t"
              << total_walks << "\t" << walks_per_sec << "\t" << visit_count << "\t" << checksum
              << "\n";
  } catch (const std::exception& e) {
    std::cerr << argv[0] << ": " << e.what() << "\n";
    return 1;
  }
  return 0;
}
