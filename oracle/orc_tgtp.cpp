// orc_tgtp.cpp -- CPU restatement of TGT+ (tau bounds, push diffusion,
// top-gamma, chi, walk counts, symmetrization) and of the graph / eigenpair
// loaders.  Test infrastructure only; see oracle.h.
//
// Every floating-point expression below keeps the reference's operand types
// (float vs double) and evaluation order.  Those expressions have to be
// alike for bit parity; the surrounding code is this oracle's own.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <random>
#include <sstream>
#include <string>
#include <vector>

float orc_walk2_mt(const uint32_t* idx, const uint32_t* adj, uint32_t s,
                   uint32_t v, uint64_t nr, uint64_t len, const float* pvec,
                   std::mt19937& gen);

namespace {

thread_local std::string g_err;

constexpr int kRwRatio = 25;  // RWRATIO, aesc.h:17
constexpr int kLanes = 8;     // BLOCKSIZE, aesc.h:446

struct Csr {
  uint64_t n, m;
  const uint32_t* idx;
  const uint32_t* adj;
  uint32_t deg(uint64_t v) const { return idx[v + 1] - idx[v]; }
};

// get_tau, aesc.h:129-135.  The reference names its first two parameters
// (dv, du) and is called as get_tau(deg(u), deg(v), ...); the formula below
// is written with those bindings made explicit.
int tau_bound(uint64_t deg_u, uint64_t deg_v, float eps, double lam, float del,
              float ups) {
  const double du = static_cast<double>(deg_v);  // reference's `du`
  const double dv = static_cast<double>(deg_u);  // reference's `dv`
  const float eps_del = eps - del > 0.0 ? eps - del : eps;
  const double a = std::log(std::max(
      (1.0 / du + 1.0 / dv - 2.0 / du / dv - ups) / eps_del / (1.0 - (lam * lam)), 1.0));
  const double b = std::log(1.0 / std::fabs(lam));
  const int tau = static_cast<int>(std::max(std::ceil(a / b - 1), 1.0));
  return tau % 2 == 0 ? tau + 1 : tau;
}

// get_upsilon, aesc.h:151-162.
float upsilon(const float* X, int omega, const double* lam, uint64_t u,
              uint64_t v, uint64_t m) {
  float ups = 0;
  const float* xu = X + u * omega;
  const float* xv = X + v * omega;
  for (int i = 1; i < omega - 1; ++i) {
    const float diff = xu[i] - xv[i];
    ups += (diff * diff) * (1 + lam[i]);
  }
  return ups / m;
}

// get_delta, aesc.h:137-150.
float spectral_delta(const float* X, int omega, const double* lam, uint64_t u,
                     uint64_t v, const double* pow_t, uint64_t m) {
  float del = 0;
  const float* xu = X + u * omega;
  const float* xv = X + v * omega;
  for (int i = 1; i < omega - 1; ++i) {
    const float diff = xu[i] - xv[i];
    const float p = static_cast<float>(pow_t[i]);
    del += (diff * diff) * p / (1 - lam[i]);
  }
  return del / m;
}

int32_t edge_tau(const Csr& g, const float* X, int omega, const double* lam,
                 const double* lam_sq, double* pow_t, float eps, uint64_t u,
                 uint64_t v) {
  const float ups = upsilon(X, omega, lam, u, v, g.m);
  const uint64_t du = g.deg(u), dv = g.deg(v);
  int tau = tau_bound(du, dv, eps, lam[1], 0, 0);  // aesc.h:188
  int t = 1;
  std::copy(lam_sq, lam_sq + omega, pow_t);
  for (;;) {  // aesc.h:191-204
    const float del = spectral_delta(X, omega, lam, u, v, pow_t, g.m);
    const int tp = tau_bound(du, dv, eps, lam[omega - 1], del, ups);
    if (t <= tp && tp < tau) {
      tau = tp;
      t += 2;
    } else {
      break;
    }
    for (int i = 0; i < omega; ++i) pow_t[i] *= lam_sq[i];
  }
  return tau;
}

// calculate_taus, aesc.h:164-216.
void all_taus(const Csr& g, const float* X, int omega, const double* lam,
              float eps, int32_t* taus, int threads) {
  std::fill(taus, taus + g.m, 0);
  std::vector<double> lam_sq(omega);
  for (int i = 0; i < omega; ++i) lam_sq[i] = lam[i] * lam[i];
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
  {
    std::vector<double> pow_t(omega);
#pragma omp for schedule(dynamic, 100)
    for (int64_t u = 0; u < static_cast<int64_t>(g.n); ++u) {
      for (uint64_t p = g.idx[u]; p < g.idx[u + 1]; ++p) {
        const uint64_t v = g.adj[p];
        if (v < static_cast<uint64_t>(u)) continue;
        const int32_t tau = edge_tau(g, X, omega, lam, lam_sq.data(), pow_t.data(), eps, u, v);
        taus[p] = tau;
        // position of the first u in v's row (aesc.h:207-209)
        const uint32_t* first = std::lower_bound(g.adj + g.idx[v], g.adj + g.idx[v + 1],
                                                 static_cast<uint32_t>(u));
        taus[first - g.adj] = tau;
      }
    }
  }
}

struct SourceScratch {
  std::vector<float> pvec;
  std::vector<char> seen;
  std::vector<char> in_top;
  std::vector<size_t> q, q_next;
  explicit SourceScratch(uint64_t n) : pvec(n), seen(n, 0), in_top(n, 0) {
    q.reserve(n);
    q_next.reserve(n);
  }
};

// get_edge_max, aesc.h:414-442.
float top_gamma_edge_max(const Csr& g, std::vector<size_t>& q, const float* pvec,
                         float gmax, int gamma, std::vector<char>& in_top) {
  const int nnz = static_cast<int>(q.size());
  const int rg = std::min(gamma, nnz);
  std::nth_element(q.begin(), q.begin() + (rg - 1), q.end(),
                   [pvec](const size_t a, const size_t b) { return pvec[a] > pvec[b]; });
  float gamma_max = 1;
  for (int i = 0; i < rg; ++i) {
    gamma_max = std::min(gamma_max, pvec[q[i]]);
    in_top[q[i]] = 1;
  }
  gamma_max = gamma < nnz ? gamma_max : 0;
  float emax = gmax + gamma_max;
  for (int i = 0; i < rg; ++i) {
    const size_t u = q[i];
    for (uint64_t p = g.idx[u]; p < g.idx[u + 1]; ++p) {
      const uint32_t w = g.adj[p];
      if (!in_top[w]) continue;
      emax = std::max(emax, pvec[u] + pvec[w]);
    }
  }
  for (int i = 0; i < rg; ++i) in_top[q[i]] = 0;
  return emax;
}

// get_chi, aesc.h:392-412.
float chi_bound(const Csr& g, uint64_t s, uint64_t v, const float* pvec,
                int len_walk, float global_min, float global_max, float edge_max) {
  float smax = 0, smin = FLT_MAX;
  for (uint64_t p = g.idx[s]; p < g.idx[s + 1]; ++p) {
    const float val = pvec[g.adj[p]];
    smax = std::max(smax, val);
    smin = std::min(smin, val);
  }
  float tmax = 0, tmin = FLT_MAX;
  for (uint64_t p = g.idx[v]; p < g.idx[v + 1]; ++p) {
    const float val = pvec[g.adj[p]];
    tmax = std::max(tmax, val);
    tmin = std::min(tmin, val);
  }
  const float chi = global_max + (smax + tmax) / 2.0 + (len_walk - 1) * edge_max -
                    smin - tmin - 2 * (len_walk - 1) * global_min;
  return chi;
}

struct Params {
  double eps, delta, logmd;
  int gamma;
  int walk_mode;
};

// One source of tgtp_baseline, aesc.h:457-526 (writes only row s of preds).
void run_source(const Csr& g, const int32_t* taus, const Params& P, uint64_t s,
                float* preds, SourceScratch& S, std::mt19937* gen,
                orc_tgtp_trace* tr) {
  float* pvec = S.pvec.data();
  std::fill(S.pvec.begin(), S.pvec.end(), 0.0f);
  int ell = 0;
  pvec[s] = 1.0;
  const float ds = static_cast<float>(g.deg(s));
  S.q.clear();
  S.q_next.clear();
  S.q.push_back(s);
  const uint64_t r0 = g.idx[s], r1 = g.idx[s + 1];
  for (uint64_t pos = r0; pos < r1; ++pos) preds[pos] = 1.0 / ds;
  float gmin, gmax;
  for (;;) {
    gmin = 1;
    gmax = 0;
    double picost = 0, rwcost = 0;
    // diffuse, aesc.h:219-239: in-place push in frontier order
    for (const size_t v : S.q) {
      const float residue = pvec[v];
      pvec[v] = 0;
      for (uint64_t p = g.idx[v]; p < g.idx[v + 1]; ++p) {
        const uint32_t u = g.adj[p];
        if (!S.seen[u]) {
          S.q_next.push_back(u);
          S.seen[u] = 1;
        }
        pvec[u] += residue / static_cast<float>(g.deg(u));
      }
    }
    for (uint64_t pos = r0; pos < r1; ++pos) preds[pos] += (pvec[s] - pvec[g.adj[pos]]) / ds;
    for (const size_t u : S.q_next) S.seen[u] = 0;
    std::swap(S.q, S.q_next);
    S.q_next.clear();
    ++ell;
    for (const size_t v : S.q) {  // aesc.h:484-491
      picost += static_cast<double>(g.deg(v));
      gmax = std::max(gmax, pvec[v]);
      gmin = std::min(gmin, pvec[v]);
    }
    gmin = S.q.size() < g.n ? 0 : gmin;
    for (uint64_t pos = r0; pos < r1; ++pos) {  // aesc.h:493-500
      const int rest = taus[pos] - ell;
      if (rest <= 0) continue;
      const double chi = 2.0 * rest * (static_cast<double>(gmax) - gmin);
      const double nr = std::max(std::ceil(8 * std::pow(chi, 2) * P.logmd / std::pow(ds * P.eps, 2)), 1.0);
      rwcost += nr;
    }
    if (picost >= kRwRatio * rwcost) break;
  }
  float emax = top_gamma_edge_max(g, S.q, pvec, gmax, P.gamma, S.in_top);
  if (P.gamma == 1) emax = 2 * gmax;
  if (tr) {
    if (tr->ell) tr->ell[s] = ell;
    if (tr->edge_max) tr->edge_max[s] = emax;
    if (tr->gmax) tr->gmax[s] = gmax;
    if (tr->gmin) tr->gmin[s] = gmin;
    if (tr->q_size) tr->q_size[s] = static_cast<uint32_t>(S.q.size());
  }
  for (uint64_t pos = r0; pos < r1; ++pos) {  // aesc.h:508-526
    const uint32_t v = g.adj[pos];
    const int64_t rest = taus[pos] - ell;
    if (tr && tr->nr) tr->nr[pos] = 0;
    if (tr && tr->walk_sum) tr->walk_sum[pos] = 0;
    if (rest <= 0) continue;
    float chi = chi_bound(g, s, v, pvec, static_cast<int>(rest), gmin, gmax, emax);
    const float dse = ds * P.eps;
    size_t nr = std::max(std::ceil(8 * (chi * chi) * P.logmd / (dse * dse)), 1.0);
    nr += (kLanes - (nr % kLanes));
    float x;
    if (P.walk_mode == ORC_WALK_MT19937)
      x = orc_walk2_mt(g.idx, g.adj, static_cast<uint32_t>(s), v, nr, rest, pvec, *gen);
    else
      x = orc_walk2_saba(g.idx, g.adj, static_cast<uint32_t>(s), v, nr,
                         static_cast<uint32_t>(rest), pvec);
    preds[pos] += x / (ds * nr);
    if (tr && tr->nr) tr->nr[pos] = nr;
    if (tr && tr->walk_sum) tr->walk_sum[pos] = x;
  }
}

// Symmetrization, aesc.h:532-550.
void symmetrize(const Csr& g, float* preds) {
  std::vector<float> rev(g.m, 0.0f);
  for (uint64_t u = 0; u < g.n; ++u) {
    for (uint64_t pos = g.idx[u]; pos < g.idx[u + 1]; ++pos) {
      const uint32_t v = g.adj[pos];
      const uint32_t* vu = std::lower_bound(g.adj + g.idx[v], g.adj + g.idx[v + 1],
                                            static_cast<uint32_t>(u));
      rev[vu - g.adj] = preds[pos];
    }
  }
  for (uint64_t i = 0; i < g.m; ++i) preds[i] += rev[i];
}

}  // namespace

extern "C" int orc_taus(uint64_t n, uint64_t m, const uint32_t* idx,
                        const uint32_t* adj, int omega, const double* lam,
                        const float* X, float eps_half, int32_t* taus,
                        int threads) {
  if (omega < 2) return -1;
  const Csr g{n, m, idx, adj};
  all_taus(g, X, omega, lam, eps_half, taus, threads);
  return 0;
}

extern "C" int orc_tgtp(uint64_t n, uint64_t m, const uint32_t* idx,
                        const uint32_t* adj, int omega, const double* lam,
                        const float* X, double epsilon, double delta, int gamma,
                        int walk_mode, int threads, float* scores,
                        orc_tgtp_trace* trace) {
  if (gamma < 1 || omega < 2) {
    g_err = "require gamma >= 1 and omega >= 2";
    return -1;
  }
  const Csr g{n, m, idx, adj};
  std::vector<int32_t> taus_local;
  int32_t* taus = trace && trace->taus ? trace->taus : nullptr;
  if (!taus) {
    taus_local.resize(m);
    taus = taus_local.data();
  }
  all_taus(g, X, omega, lam, static_cast<float>(epsilon / 2.0), taus, threads);  // aesc.h:122
  const Params P{epsilon, delta, std::log(m / delta), gamma, walk_mode};
  std::fill(scores, scores + m, 0.0f);  // aesc.h:445
  if (walk_mode == ORC_WALK_MT19937) {
    std::mt19937 gen(42);  // aesc_baseline.h:17,261
    SourceScratch S(n);
    for (uint64_t s = 0; s < n; ++s) run_source(g, taus, P, s, scores, S, &gen, trace);
  } else {
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
      SourceScratch S(n);
#pragma omp for schedule(dynamic)
      for (int64_t s = 0; s < static_cast<int64_t>(n); ++s)
        run_source(g, taus, P, s, scores, S, nullptr, trace);
    }
  }
  if (trace && trace->preds_row) std::copy(scores, scores + m, trace->preds_row);
  symmetrize(g, scores);
  return 0;
}

extern "C" int orc_tgtp_rows(uint64_t n, uint64_t m, const uint32_t* idx,
                             const uint32_t* adj, const int32_t* taus,
                             double epsilon, double delta, int gamma,
                             uint32_t src_begin, uint32_t src_end, int threads,
                             float* preds_rows) {
  if (gamma < 1) return -1;
  const Csr g{n, m, idx, adj};
  const Params P{epsilon, delta, std::log(m / delta), gamma, ORC_WALK_SABA};
  if (src_end > n) src_end = static_cast<uint32_t>(n);
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
  {
    SourceScratch S(n);
#pragma omp for schedule(dynamic)
    for (int64_t s = src_begin; s < static_cast<int64_t>(src_end); ++s)
      run_source(g, taus, P, s, preds_rows, S, nullptr, nullptr);
  }
  return 0;
}

extern "C" int orc_tgtp_sources(uint64_t n, uint64_t m, const uint32_t* idx,
                                const uint32_t* adj, const int32_t* taus, double epsilon,
                                double delta, int gamma, const uint32_t* sources,
                                uint64_t count, int threads, float* preds_rows) {
  if (gamma < 1) return -1;
  for (uint64_t i = 0; i < count; ++i)
    if (sources[i] >= n) return -1;
  const Csr g{n, m, idx, adj};
  const Params P{epsilon, delta, std::log(m / delta), gamma, ORC_WALK_SABA};
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
  {
    SourceScratch S(n);
#pragma omp for schedule(dynamic)
    for (int64_t i = 0; i < static_cast<int64_t>(count); ++i)
      run_source(g, taus, P, sources[i], preds_rows, S, nullptr, nullptr);
  }
  return 0;
}

// graph_t::load_txt, graph.h:29-97 (undirected, no skipped lines).
extern "C" int orc_load_txt(const char* path, uint64_t* n_out, uint64_t* m_out,
                            uint32_t** idx_out, uint32_t** adj_out) {
  std::ifstream in(path);
  if (!in) {
    g_err = std::string("cannot open graph: ") + path;
    return -1;
  }
  std::vector<std::vector<uint32_t>> rows;
  uint64_t n = 0;
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && (line[0] == '#' || line[0] == '%')) continue;
    std::stringstream ss(line);
    ss >> std::ws;
    if (ss.eof() || ss.peek() == '#' || ss.peek() == '%') continue;
    int64_t a, b;
    if (!(ss >> a >> b) || a < 0 || b < 0 || a >= 0xFFFFFFFFll || b >= 0xFFFFFFFFll) {
      g_err = std::string("invalid edge in ") + path + ": " + line;
      return -1;
    }
    const uint32_t s = static_cast<uint32_t>(a), t = static_cast<uint32_t>(b);
    const uint64_t hi = static_cast<uint64_t>(std::max(s, t)) + 1;
    if (rows.size() < hi) rows.resize(hi);
    rows[s].push_back(t);
    rows[t].push_back(s);
    n = std::max(hi, n);
  }
  if (rows.empty()) {
    g_err = std::string("empty graph: ") + path;
    return -1;
  }
  uint64_t m = 0;
  for (auto& r : rows) m += r.size();
  if (m > 0xFFFFFFFFull) {
    g_err = "graph exceeds CSR index capacity";
    return -1;
  }
  for (auto& r : rows) std::sort(r.begin(), r.end());
  uint32_t* idx = static_cast<uint32_t*>(std::malloc((n + 1) * sizeof(uint32_t)));
  uint32_t* adj = static_cast<uint32_t*>(std::malloc((m ? m : 1) * sizeof(uint32_t)));
  idx[0] = 0;
  uint64_t pos = 0;
  for (uint64_t v = 0; v < n; ++v) {
    for (uint32_t w : rows[v]) adj[pos++] = w;
    idx[v + 1] = static_cast<uint32_t>(pos);
  }
  *n_out = n;
  *m_out = m;
  *idx_out = idx;
  *adj_out = adj;
  return 0;
}

// EigenVecs::load_txt, aesc.h:51-64: eigenvector entries are read straight
// into float (not double then float).
extern "C" int orc_load_eigens(const char* path, uint64_t n, int omega,
                               double* lam, float* X) {
  std::ifstream in(path);
  if (!in) {
    g_err = std::string("cannot open eigenpairs: ") + path;
    return -1;
  }
  for (int i = 0; i < omega; ++i) {
    if (!(in >> lam[i]) || !std::isfinite(lam[i])) {
      g_err = std::string("invalid eigenvalue in ") + path;
      return -1;
    }
    for (uint64_t j = 0; j < n; ++j) {
      float& x = X[static_cast<uint64_t>(omega) * j + i];
      if (!(in >> x) || !std::isfinite(x)) {
        g_err = std::string("invalid eigenvector in ") + path;
        return -1;
      }
    }
  }
  return 0;
}

extern "C" void orc_free(void* p) { std::free(p); }
extern "C" const char* orc_last_error(void) { return g_err.c_str(); }
