// host_tgtp.cpp -- host-side pieces of the tgtp pipeline that must use the
// C library's transcendental functions: the tau of an edge whose device
// result sat within rounding distance of an integer boundary of ceil(),
// plus the tables the device kernels read.  Compiled with g++,
// -ffp-contract=off, so every float/double expression evaluates as written.
#include "host_tgtp.hpp"

#include <algorithm>
#include <cmath>

namespace tgtp_host {

// get_tau, aesc.h:129-135 (called as get_tau(deg(u), deg(v), ...); the
// reference's parameter names are swapped relative to the call).
int tau_bound(uint64_t deg_u, uint64_t deg_v, float eps, double lam, float del, float ups) {
  const double du = static_cast<double>(deg_v);
  const double dv = static_cast<double>(deg_u);
  const float eps_del = eps - del > 0.0 ? eps - del : eps;
  const double a = std::log(std::max(
      (1.0 / du + 1.0 / dv - 2.0 / du / dv - ups) / eps_del / (1.0 - (lam * lam)), 1.0));
  const double b = std::log(1.0 / std::fabs(lam));
  const int tau = static_cast<int>(std::max(std::ceil(a / b - 1), 1.0));
  return tau % 2 == 0 ? tau + 1 : tau;
}

// calculate_taus' per-edge body, aesc.h:182-204.
int edge_tau(const uint32_t* idx, const float* X, int omega, const double* lam, uint64_t m,
             float eps, uint32_t u, uint32_t v) {
  const float* xu = X + static_cast<uint64_t>(u) * omega;
  const float* xv = X + static_cast<uint64_t>(v) * omega;
  float ups = 0;
  for (int i = 1; i < omega - 1; ++i) {
    const float diff = xu[i] - xv[i];
    ups += (diff * diff) * (1 + lam[i]);
  }
  ups = ups / m;
  const uint64_t du = idx[u + 1] - idx[u], dv = idx[v + 1] - idx[v];
  int tau = tau_bound(du, dv, eps, lam[1], 0, 0);
  int t = 1;
  std::vector<double> sq(omega), pw(omega);
  for (int i = 0; i < omega; ++i) sq[i] = lam[i] * lam[i];
  pw = sq;
  for (;;) {
    float del = 0;
    for (int i = 1; i < omega - 1; ++i) {
      const float diff = xu[i] - xv[i];
      const float p = static_cast<float>(pw[i]);
      del += (diff * diff) * p / (1 - lam[i]);
    }
    del = del / m;
    const int tp = tau_bound(du, dv, eps, lam[omega - 1], del, ups);
    if (t <= tp && tp < tau) {
      tau = tp;
      t += 2;
    } else {
      break;
    }
    for (int i = 0; i < omega; ++i) pw[i] *= sq[i];
  }
  return tau;
}

std::vector<double> power_table(const double* lam, int omega, uint32_t k_rows) {
  // row k = lambda_i^(2(k+1)) built by repeated products, as aesc.h:190,201-203
  std::vector<double> t(static_cast<size_t>(k_rows) * omega);
  for (int i = 0; i < omega; ++i) {
    const double sq = lam[i] * lam[i];
    double p = sq;
    for (uint32_t k = 0; k < k_rows; ++k) {
      t[static_cast<size_t>(k) * omega + i] = p;
      p *= sq;
    }
  }
  return t;
}

int symmetrize_sources(const uint32_t* idx, const uint32_t* adj, uint64_t n, uint64_t m,
                       std::vector<int32_t>& src) {
  const int64_t rows = static_cast<int64_t>(n);
  // host threads only for large graphs: on small ones waking them (and their
  // spinning afterwards) costs more than the loop, and delays the launches
  // that follow
  const bool par = m >= (1ull << 20);
  int unsorted = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : unsorted) if (par)
  for (int64_t u = 0; u < rows; ++u)
    for (uint64_t p = idx[u] + 1; p < idx[u + 1]; ++p)
      if (adj[p - 1] > adj[p]) {
        ++unsorted;
        break;
      }
  if (unsorted) return 1;
  // aesc.h:538-545: rev_preds[first position of u in row v] = preds[pos(u,v)],
  // rows u ascending, positions ascending (a later duplicate overwrites).
  // Slot (first u in row v) is written only from row u, so rows run in
  // parallel and each row in position order: the same map.
  src.assign(m, -1);
  int missing = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : missing) if (par)
  for (int64_t u = 0; u < rows; ++u) {
    for (uint64_t p = idx[u]; p < idx[u + 1]; ++p) {
      const uint32_t v = adj[p];
      const uint32_t* end = adj + idx[v + 1];
      const uint32_t* first = std::lower_bound(adj + idx[v], end, static_cast<uint32_t>(u));
      if (first == end || *first != u) {
        ++missing;
        break;
      }
      src[first - adj] = static_cast<int32_t>(p);
    }
  }
  return missing ? 2 : 0;
}

}  // namespace tgtp_host
