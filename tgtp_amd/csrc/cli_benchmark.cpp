// benchmark drop-in (src/benchmark.cpp:9-53): same arguments, header and
// TSV rows at setprecision(10); each measured run times one tgtp() call on
// the GPU with the graph and eigenpairs already resident in HBM (loading
// excluded, as benchmark.cpp:40-42).  The threads column echoes the CSV item:
// the GPU path has no host thread count.
//
// usage: benchmark dataset [eps_csv=0.05,0.01,0.005] [threads_csv=1]
//                  [repeats=3] [data_root=datasets] [omega=128] [gamma=10] [warmups=1]
//
// Built with TGTP_BENCH_BASELINE=1 it is the benchmark_baseline drop-in
// (src/benchmark_baseline.cpp): the reference's baseline tgtp with the
// MT19937 2-way walks (TGTP_MODE_FAITHFUL | TGTP_WALK2_MT19937), threads
// must be 1, rows labelled mc2way-baseline.
#include <chrono>
#include <cmath>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <vector>

#include "cli_common.hpp"

#ifndef TGTP_BENCH_BASELINE
#define TGTP_BENCH_BASELINE 0
#endif

int main(int argc, char** argv) {
  if (argc < 2 || argc > 9) {
    if (TGTP_BENCH_BASELINE)
      std::cerr << "Usage: benchmark_baseline dataset [eps_csv=0.05,0.01,0.005] [threads=1]"
                << " [repeats=3] [data_root=datasets] [omega=128] [gamma=10] [warmups=1]\n";
    else
      std::cerr << "Usage: benchmark dataset [eps_csv=0.05,0.01,0.005] [threads_csv=1]"
                << " [repeats=3] [data_root=datasets] [omega=128] [gamma=10] [warmups=1]\n";
    return 1;
  }
  try {
    const auto eps = cli::split_csv(argc > 2 ? argv[2] : "0.05,0.01,0.005");
    const auto threads = cli::split_csv(argc > 3 ? argv[3] : "1");
    if (TGTP_BENCH_BASELINE && (threads.size() != 1 || cli::positive(threads[0]) != 1))
      throw std::invalid_argument("baseline requires threads=1");
    const uint32_t repeats = cli::positive(argc > 4 ? argv[4] : "3");
    const double warmup_arg = cli::number(argc > 8 ? argv[8] : "1");
    if (warmup_arg < 0 || warmup_arg > INT32_MAX || warmup_arg != std::floor(warmup_arg))
      throw std::invalid_argument("invalid warmups");
    const unsigned warmups = static_cast<unsigned>(warmup_arg);
    const std::string root = argc > 5 ? argv[5] : "datasets";
    const std::string dataset = argv[1];
    const int32_t omega = static_cast<int32_t>(cli::positive(argc > 6 ? argv[6] : "128"));
    const int32_t gamma = static_cast<int32_t>(cli::positive(argc > 7 ? argv[7] : "10"));
    for (const auto& e : eps)
      if (!(cli::number(e) > 0 && cli::number(e) < 1))
        throw std::invalid_argument("epsilon must be between zero and one");
    for (const auto& t : threads) cli::positive(t);
    const auto dir = std::filesystem::path(root) / dataset;
    cli::Csr csr;
    cli::load_graph((dir / "graph.txt").string(), csr);
    char err[512] = {0};
    if (tgtp_check_aesc_graph(&csr.g, omega, err, sizeof(err))) throw std::runtime_error(err);
    std::vector<double> lam(omega);
    std::vector<float> X(csr.g.n * static_cast<uint64_t>(omega));
    const std::string epath = (dir / ("sorted_eigens_" + std::to_string(omega) + ".txt")).string();
    cli::load_eigens(epath, csr.g.n, omega, lam.data(), X.data());
    cli::Ctx ctx(cli::device_from_env());
    const tgtp_graph gv = csr.view();
    tgtp_dgraph* dg = nullptr;
    ctx.check(tgtp_graph_upload(ctx.c, &gv, &dg), "tgtp_graph_upload");
    const tgtp_eigens ev{omega, lam.data(), X.data()};
    const int32_t mode =
        TGTP_BENCH_BASELINE ? (TGTP_MODE_FAITHFUL | TGTP_WALK2_MT19937) : cli::mode_from_env();
    const char* sampler = TGTP_BENCH_BASELINE ? "mc2way-baseline" : "seed32-grid-multiply-high";
    std::vector<float> scores(csr.g.m);
    std::cout << std::setprecision(10) << std::unitbuf;
    std::cout << "sampler\tdataset\tepsilon\tthreads\trepeat\tseconds\tchecksum\n";
    for (const auto& e : eps) {
      for (const auto& t : threads) {
        const tgtp_params pv{cli::number(e), 1.0 / static_cast<double>(csr.g.n), gamma, mode};
        for (uint64_t run = 0; run < uint64_t(warmups) + repeats; ++run) {
          const auto start = std::chrono::steady_clock::now();
          ctx.check(tgtp_score_dev(ctx.c, dg, &ev, &pv, scores.data(), nullptr), "tgtp_score_dev");
          const double elapsed =
              std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
          double checksum = 0;
          for (size_t j = 0; j < csr.g.m; ++j) {
            if (!std::isfinite(scores[j])) throw std::runtime_error("nonfinite score");
            checksum += scores[j];
          }
          if (run >= warmups)
            std::cout << sampler << '\t' << dataset << '\t' << pv.epsilon << '\t' << t << '\t'
                      << run - warmups << '\t' << elapsed << '\t' << checksum << '\n';
        }
      }
    }
    tgtp_graph_free(ctx.c, dg);
  } catch (const std::exception& error) {
    std::cerr << "benchmark: " << error.what() << '\n';
    return 1;
  }
  return 0;
}
