// tgtp_aesc.hpp -- header-only C++ drop-in for the reference's AESC
// interface (src/aesc.h:83-127): AESCGraph, AESCConfig and AESC::tgtp()
// with the same names, members and ownership, computed on the GPU through
// the C ABI of tgtp_hip.h.  A caller written against aesc.h
// (src/main.cpp:34-43, src/benchmark.cpp:29-41) compiles unchanged after
// replacing `#include "aesc.h"` with `#include "tgtp_aesc.hpp"` and linking
// libtgtp_hip.so.
//
//   * AESCGraph(graph_path, eigen_path, omega): loads graph.txt and
//     sorted_eigens_<omega>.txt with the reference's semantics and throws
//     std::runtime_error on the same conditions (graph.h:34-74, aesc.h:53-61,
//     87-89).  Members n, m, idx, adj, degree(), begin(), end(), eigens
//     (values / vectors / omega) as in graph_t / EigenVecs<float>.
//   * AESC(g, config).tgtp() returns `new float[g.m]` owned by the caller
//     (aesc.h:120-127); errors throw std::runtime_error.
//   * The GPU context is created on first use, one per thread (device
//     TGTP_DEVICE, default 0); TGTP_MODE=faithful|fp32|bf16 selects the
//     arithmetic (default faithful = bit-exact with the reference).
#pragma once

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tgtp_graph_io.h"
#include "tgtp_hip.h"

namespace tgtp_dropin {

inline tgtp_ctx* thread_context() {
  struct Holder {
    tgtp_ctx* c = nullptr;
    ~Holder() { tgtp_ctx_destroy(c); }
  };
  thread_local Holder h;
  if (!h.c) {
    const char* d = std::getenv("TGTP_DEVICE");
    const int rc = tgtp_ctx_create(d && *d ? std::atoi(d) : 0, &h.c);
    if (rc) throw std::runtime_error(std::string("tgtp: cannot open HIP device: ") + tgtp_strerror(rc));
  }
  return h.c;
}

inline int32_t mode_from_env() {
  const char* s = std::getenv("TGTP_MODE");
  const std::string m = s ? s : "";
  if (m.empty() || m == "faithful") return TGTP_MODE_FAITHFUL;
  if (m == "fp32") return TGTP_MODE_FAST_FP32;
  if (m == "bf16") return TGTP_MODE_FAST_BF16;
  throw std::invalid_argument("TGTP_MODE must be faithful, fp32 or bf16");
}

}  // namespace tgtp_dropin

template <typename T>
struct EigenVecs {
  std::vector<double> value_store;
  std::vector<T> vector_store;
  double* values = nullptr;
  T* vectors = nullptr;
  size_t len = 0, omega = 0;
  T* begin(size_t i) { return &vectors[i * omega]; }
  double& operator[](size_t i) { return values[i]; }
};

struct AESCGraph {
  uint32_t* idx = nullptr;
  uint32_t* adj = nullptr;
  size_t n = 0, m = 0;
  EigenVecs<float> eigens;

  AESCGraph(const std::string& graph_path, const std::string& eigen_path, size_t omega) {
    char err[512] = {0};
    int rc = tgtp_load_txt(graph_path.c_str(), &csr_.c, err, sizeof(err));
    if (rc) throw std::runtime_error(err);
    n = csr_.c.n;
    m = csr_.c.m;
    idx = csr_.c.idx;
    adj = csr_.c.adj;
    if (tgtp_check_aesc_graph(&csr_.c, static_cast<int32_t>(omega), err, sizeof(err)))
      throw std::runtime_error(err);
    eigens.len = n;
    eigens.omega = omega;
    eigens.value_store.resize(omega);
    eigens.vector_store.resize(n * omega);
    eigens.values = eigens.value_store.data();
    eigens.vectors = eigens.vector_store.data();
    rc = tgtp_load_eigens_txt(eigen_path.c_str(), n, static_cast<int32_t>(omega), eigens.values,
                              eigens.vectors, err, sizeof(err));
    if (rc) throw std::runtime_error(err);
  }
  AESCGraph(const AESCGraph&) = delete;
  AESCGraph& operator=(const AESCGraph&) = delete;
  uint32_t degree(uint32_t v) const { return idx[v + 1] - idx[v]; }
  uint32_t* begin(uint32_t v) { return adj + idx[v]; }
  uint32_t* end(uint32_t v) { return adj + idx[v + 1]; }
  void free() {}  // storage is released by the destructor

 private:
  // owns the loaded CSR; a member, so that it is also released when the
  // constructor throws after loading the graph
  struct Owned {
    tgtp_csr c{};
    Owned() = default;
    Owned(const Owned&) = delete;
    Owned& operator=(const Owned&) = delete;
    ~Owned() { tgtp_csr_free(&c); }
  } csr_;
};

struct AESCConfig {
  double epsilon, delta;
  std::string strFolder, strGraph;
  int omega, gamma;
};

struct AESC {
  AESCGraph& g;
  AESCConfig& config;
  AESC(AESCGraph& g, AESCConfig& config) : g(g), config(config) {}

  // aesc.h:120-127: scores in CSR order, caller owns (delete[]).
  float* tgtp() {
    tgtp_ctx* ctx = tgtp_dropin::thread_context();
    const tgtp_graph gv{g.n, g.m, g.idx, g.adj};
    const tgtp_eigens ev{static_cast<int32_t>(g.eigens.omega), g.eigens.values, g.eigens.vectors};
    const tgtp_params pv{config.epsilon, config.delta, config.gamma, tgtp_dropin::mode_from_env()};
    std::unique_ptr<float[]> out(new float[g.m]);
    const int rc = tgtp_score(ctx, &gv, &ev, &pv, out.get(), nullptr);
    if (rc)
      throw std::runtime_error(std::string("tgtp: ") + tgtp_strerror(rc) + ": " +
                               tgtp_ctx_last_error(ctx));
    return out.release();
  }
};
