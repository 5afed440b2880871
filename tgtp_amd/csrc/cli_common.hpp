// cli_common.hpp -- argument parsing shared by the command-line drop-ins,
// with the reference's validation rules (main.cpp:18-30,
// benchmark_options.h:9-30).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
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

inline void load_graph(const std::string& path, Csr& out) {
  char err[512] = {0};
  const int rc = tgtp_load_txt(path.c_str(), &out.g, err, sizeof(err));
  if (rc) throw std::runtime_error(err);
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
