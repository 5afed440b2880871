// cli_common.hpp -- argument parsing shared by the command-line drop-ins,
// with the reference's validation rules (main.cpp:18-30,
// benchmark_options.h:9-30).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <filesystem>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "tgtp_graph_io.h"
#include "tgtp_hip.h"

namespace cli {

inline std::vector<std::string> split_csv(const std::string& text) {
  std::vector<std::string> out;
  std::stringstream in(text);
  std::string tok;
  while (std::getline(in, tok, ',')) {
    if (tok.empty()) throw std::invalid_argument("empty CSV item");
    out.push_back(tok);
  }
  if (out.empty() || text.back() == ',') throw std::invalid_argument("empty CSV list/item");
  return out;
}

inline double number(const std::string& text) {
  size_t used = 0;
  const double v = std::stod(text, &used);
  if (used != text.size() || !std::isfinite(v))
    throw std::invalid_argument("invalid number: " + text);
  return v;
}

inline uint32_t positive(const std::string& text) {
  const double v = number(text);
  if (v < 1 || v > INT32_MAX || v != std::floor(v))
    throw std::invalid_argument("invalid positive integer: " + text);
  return static_cast<uint32_t>(v);
}

// TGTP_MODE=faithful|fp32|bf16 selects the arithmetic (default faithful)
inline int32_t mode_from_env() {
  const char* s = std::getenv("TGTP_MODE");
  if (!s || !*s) return TGTP_MODE_FAITHFUL;
  const std::string m(s);
  if (m == "faithful") return TGTP_MODE_FAITHFUL;
  if (m == "fp32") return TGTP_MODE_FAST_FP32;
  if (m == "bf16") return TGTP_MODE_FAST_BF16;
  throw std::invalid_argument("TGTP_MODE must be faithful, fp32 or bf16");
}

struct Csr {
  tgtp_csr g{};
  ~Csr() { tgtp_csr_free(&g); }
  tgtp_graph view() const { return tgtp_graph{g.n, g.m, g.idx, g.adj}; }
};

// TGTP_CACHE=1: text inputs go through binary caches beside them
// (<graph.txt>.csr, <eigens.txt>.bin), used while newer than the text and
// whole, rebuilt otherwise (a cache that cannot be written is skipped).
inline bool cache_from_env() {
  const char* s = std::getenv("TGTP_CACHE");
  return s && std::string(s) == "1";
}

inline bool cache_fresh(const std::string& text, const std::string& cache) {
  std::error_code ec;
  const auto tc = std::filesystem::last_write_time(cache, ec);
  if (ec) return false;
  const auto tt = std::filesystem::last_write_time(text, ec);
  return !ec && tc >= tt;
}

inline void load_graph(const std::string& path, Csr& out) {
  char err[512] = {0};
  const std::string cache = path + ".csr";
  const bool use_cache = cache_from_env();
  if (use_cache && cache_fresh(path, cache) &&
      tgtp_load_csr_bin(cache.c_str(), &out.g, err, sizeof(err)) == 0)
    return;
  const int rc = tgtp_load_txt(path.c_str(), &out.g, err, sizeof(err));
  if (rc) throw std::runtime_error(err);
  if (use_cache) tgtp_save_csr_bin(&out.g, cache.c_str());
}

// values [omega], vectors [n * omega] vertex-major (aesc.h:51-64)
inline void load_eigens(const std::string& path, uint64_t n, int32_t omega, double* values,
                        float* vectors) {
  char err[512] = {0};
  const std::string cache = path + ".bin";
  const bool use_cache = cache_from_env();
  if (use_cache && cache_fresh(path, cache) &&
      tgtp_load_eigens_bin(cache.c_str(), n, omega, values, vectors, err, sizeof(err)) == 0)
    return;
  if (tgtp_load_eigens_txt(path.c_str(), n, omega, values, vectors, err, sizeof(err)))
    throw std::runtime_error(err);
  if (use_cache) tgtp_save_eigens_bin(cache.c_str(), n, omega, values, vectors);
}

struct Ctx {
  tgtp_ctx* c = nullptr;
  explicit Ctx(int device) {
    const int rc = tgtp_ctx_create(device, &c);
    if (rc) throw std::runtime_error(std::string("cannot open HIP device: ") + tgtp_strerror(rc));
  }
  ~Ctx() { tgtp_ctx_destroy(c); }
  void check(int rc, const char* what) const {
    if (rc)
      throw std::runtime_error(std::string(what) + ": " + tgtp_strerror(rc) + ": " +
                               tgtp_ctx_last_error(c));
  }
};

inline int device_from_env() {
  const char* s = std::getenv("TGTP_DEVICE");
  return s && *s ? std::atoi(s) : 0;
}

}  // namespace cli
