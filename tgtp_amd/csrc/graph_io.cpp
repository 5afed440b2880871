// graph_io.cpp -- host graph / eigenpair loading, seeded generators and CSR
// helpers of libtgtp_hip.so (declarations and semantics: tgtp_graph_io.h).
#include "tgtp_graph_io.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <unordered_set>
#include <utility>
#include <vector>

namespace {

void set_err(char* err, size_t errlen, const std::string& msg) {
  if (!err || errlen == 0) return;
  std::snprintf(err, errlen, "%s", msg.c_str());
}

inline bool is_space(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

// Reads an optionally signed decimal integer the way istream >> int64_t
// does on one line: leading blanks skipped, at least one digit required,
// overflow is a failure.
bool parse_i64(const char*& p, const char* end, int64_t& out) {
  while (p < end && is_space(*p)) ++p;
  if (p == end) return false;
  bool neg = false;
  if (*p == '+' || *p == '-') {
    neg = (*p == '-');
    ++p;
  }
  if (p == end || *p < '0' || *p > '9') return false;
  unsigned __int128 v = 0;
  const unsigned __int128 lim = neg ? (static_cast<unsigned __int128>(INT64_MAX) + 1) : INT64_MAX;
  while (p < end && *p >= '0' && *p <= '9') {
    v = v * 10 + static_cast<unsigned>(*p - '0');
    if (v > lim) return false;
    ++p;
  }
  out = neg ? static_cast<int64_t>(-static_cast<__int128>(v)) : static_cast<int64_t>(v);
  return true;
}

struct Mapped {
  const char* data = nullptr;
  size_t size = 0;
  int fd = -1;
  ~Mapped() {
    if (data && size) munmap(const_cast<char*>(data), size);
    if (fd >= 0) close(fd);
  }
  bool open_file(const char* path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) != 0) return false;
    size = static_cast<size_t>(st.st_size);
    if (size == 0) return true;
    void* p = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) {
      size = 0;
      return false;
    }
    data = static_cast<const char*>(p);
    return true;
  }
};

// Builds a sorted-row CSR holding both directions of each (a, b) pair.
int build_csr(uint64_t n, const std::vector<std::pair<uint32_t, uint32_t>>& edges,
              tgtp_csr* out) {
  const uint64_t m = 2ull * edges.size();
  if (m > 0xFFFFFFFFull) return -1;
  uint32_t* idx = static_cast<uint32_t*>(std::malloc((n + 1) * sizeof(uint32_t)));
  uint32_t* adj = static_cast<uint32_t*>(std::malloc((m ? m : 1) * sizeof(uint32_t)));
  if (!idx || !adj) {
    std::free(idx);
    std::free(adj);
    return -3;
  }
  std::vector<uint64_t> cnt(n + 1, 0);
  for (auto& e : edges) {
    ++cnt[e.first];
    ++cnt[e.second];
  }
  idx[0] = 0;
  for (uint64_t v = 0; v < n; ++v) idx[v + 1] = static_cast<uint32_t>(idx[v] + cnt[v]);
  std::vector<uint64_t> fill(idx, idx + n);
  for (auto& e : edges) {
    adj[fill[e.first]++] = e.second;
    adj[fill[e.second]++] = e.first;
  }
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < static_cast<int64_t>(n); ++v) std::sort(adj + idx[v], adj + idx[v + 1]);
  out->n = n;
  out->m = m;
  out->idx = idx;
  out->adj = adj;
  return 0;
}

inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
// uniform in (0, 1]
inline double unit_open0(uint64_t& s) { return static_cast<double>((splitmix64(s) >> 11) + 1) * 0x1.0p-53; }
// uniform integer in [0, k)
inline uint64_t below(uint64_t& s, uint64_t k) {
  return static_cast<uint64_t>((static_cast<unsigned __int128>(splitmix64(s)) * k) >> 64);
}
inline uint64_t stream_seed(uint64_t seed, uint64_t salt) {
  uint64_t s = seed ^ (salt * 0xd1b54a32d192ed03ull);
  return splitmix64(s);
}

struct Dsu {
  std::vector<uint32_t> p;
  explicit Dsu(uint64_t n) : p(n) { std::iota(p.begin(), p.end(), 0u); }
  uint32_t find(uint32_t x) {
    while (p[x] != x) {
      p[x] = p[p[x]];
      x = p[x];
    }
    return x;
  }
  void unite(uint32_t a, uint32_t b) {
    a = find(a);
    b = find(b);
    if (a == b) return;
    if (a < b) std::swap(a, b);
    p[a] = b;  // root = smaller id
  }
};

// Joins every component other than the largest to the largest by one edge
// from the component's smallest vertex to a weight-sampled vertex of the
// largest component.  Components are visited in order of smallest vertex.
void connect_components(uint64_t n, const std::vector<double>& w, uint64_t seed,
                        std::vector<std::pair<uint32_t, uint32_t>>& edges) {
  Dsu dsu(n);
  for (auto& e : edges) dsu.unite(e.first, e.second);
  std::vector<uint64_t> size(n, 0);
  for (uint64_t v = 0; v < n; ++v) ++size[dsu.find(static_cast<uint32_t>(v))];
  uint32_t giant = 0;
  for (uint64_t v = 0; v < n; ++v)
    if (size[v] > size[giant]) giant = static_cast<uint32_t>(v);
  std::vector<uint32_t> members;
  std::vector<double> cum;
  double acc = 0;
  for (uint64_t v = 0; v < n; ++v) {
    if (dsu.find(static_cast<uint32_t>(v)) != giant) continue;
    members.push_back(static_cast<uint32_t>(v));
    acc += w[v];
    cum.push_back(acc);
  }
  uint64_t rs = stream_seed(seed, 0xC0FFEEull);
  std::vector<char> done(n, 0);
  for (uint64_t v = 0; v < n; ++v) {
    const uint32_t r = dsu.find(static_cast<uint32_t>(v));
    if (r == giant || done[r]) continue;
    done[r] = 1;  // v is the smallest vertex of its component (roots are minima)
    const double x = unit_open0(rs) * acc;
    size_t k = static_cast<size_t>(std::lower_bound(cum.begin(), cum.end(), x) - cum.begin());
    if (k >= members.size()) k = members.size() - 1;
    edges.emplace_back(static_cast<uint32_t>(v), members[k]);
  }
}

}  // namespace

extern "C" int tgtp_load_txt(const char* path, tgtp_csr* out, char* err, size_t errlen) {
  if (!path || !out) {
    set_err(err, errlen, "null argument");
    return -1;
  }
  Mapped f;
  if (!f.open_file(path)) {
    set_err(err, errlen, std::string("cannot open graph: ") + path);
    return -4;
  }
  std::vector<std::pair<uint32_t, uint32_t>> edges;
  uint64_t n = 0;
  const char* p = f.data;
  const char* end = f.data + f.size;
  while (p < end) {
    const char* eol = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
    if (!eol) eol = end;
    const char* line = p;
    p = eol + 1;
    if (line < eol && (*line == '#' || *line == '%')) continue;
    const char* q = line;
    while (q < eol && is_space(*q)) ++q;
    if (q == eol || *q == '#' || *q == '%') continue;
    int64_t a, b;
    if (!parse_i64(q, eol, a) || !parse_i64(q, eol, b) || a < 0 || b < 0 ||
        a >= 0xFFFFFFFFll || b >= 0xFFFFFFFFll) {
      set_err(err, errlen, std::string("invalid edge in ") + path + ": " +
                               std::string(line, static_cast<size_t>(eol - line)));
      return -1;
    }
    edges.emplace_back(static_cast<uint32_t>(a), static_cast<uint32_t>(b));
    n = std::max<uint64_t>(n, static_cast<uint64_t>(std::max(a, b)) + 1);
  }
  if (edges.empty()) {
    set_err(err, errlen, std::string("empty graph: ") + path);
    return -1;
  }
  if (2ull * edges.size() > 0xFFFFFFFFull) {
    set_err(err, errlen, "graph exceeds CSR index capacity");
    return -1;
  }
  const int rc = build_csr(n, edges, out);
  if (rc == -3) set_err(err, errlen, "out of host memory");
  return rc;
}

extern "C" void tgtp_csr_free(tgtp_csr* g) {
  if (!g) return;
  std::free(g->idx);
  std::free(g->adj);
  g->idx = nullptr;
  g->adj = nullptr;
  g->n = g->m = 0;
}

extern "C" int tgtp_load_eigens_txt(const char* path, uint64_t n, int32_t omega,
                                    double* values, float* vectors, char* err,
                                    size_t errlen) {
  Mapped f;
  if (!f.open_file(path)) {
    set_err(err, errlen, std::string("cannot open eigenpairs: ") + path);
    return -4;
  }
  // strtod / strtof need a terminated buffer; copy once.
  std::string buf(f.data ? f.data : "", f.size);
  const char* p = buf.c_str();
  for (int32_t i = 0; i < omega; ++i) {
    char* next = nullptr;
    errno = 0;
    const double lam = std::strtod(p, &next);
    if (next == p || !std::isfinite(lam)) {
      set_err(err, errlen, std::string("invalid eigenvalue in ") + path);
      return -1;
    }
    values[i] = lam;
    p = next;
    for (uint64_t j = 0; j < n; ++j) {
      const float x = std::strtof(p, &next);  // parsed straight to float, as aesc.h:60
      if (next == p || !std::isfinite(x)) {
        set_err(err, errlen, std::string("invalid eigenvector in ") + path);
        return -1;
      }
      vectors[static_cast<uint64_t>(omega) * j + i] = x;
      p = next;
    }
  }
  return 0;
}

extern "C" int tgtp_check_aesc_graph(const tgtp_csr* g, int32_t omega, char* err, size_t errlen) {
  if (omega < 2 || static_cast<uint64_t>(omega) > g->n) {
    set_err(err, errlen, "omega must be between 2 and vertex count");
    return -1;
  }
  for (uint64_t v = 0; v < g->n; ++v) {
    if (g->idx[v + 1] == g->idx[v]) {
      set_err(err, errlen, "graph vertex IDs must be contiguous with no isolated vertices");
      return -1;
    }
  }
  return 0;
}

extern "C" int tgtp_gen_chung_lu(uint64_t n, double beta, double mean_deg, double w_cap,
                                 uint64_t seed, tgtp_csr* out) {
  if (n < 2 || n > 0xFFFFFFFEull || !(beta > 1.0) || !(mean_deg > 0)) return -1;
  std::vector<double> w(n);
  const double ex = -1.0 / (beta - 1.0);
  double sum = 0;
  for (uint64_t i = 0; i < n; ++i) {
    w[i] = std::pow(static_cast<double>(i + 1), ex);
    sum += w[i];
  }
  const double c = mean_deg * static_cast<double>(n) / sum;
  double S = 0;
  for (uint64_t i = 0; i < n; ++i) {
    w[i] *= c;
    if (w_cap > 0 && w[i] > w_cap) w[i] = w_cap;
    S += w[i];
  }
  constexpr uint64_t kChunk = 1024;
  const uint64_t nchunks = (n + kChunk - 1) / kChunk;
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> parts(nchunks);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t ci = 0; ci < static_cast<int64_t>(nchunks); ++ci) {
    auto& part = parts[ci];
    const uint64_t u0 = static_cast<uint64_t>(ci) * kChunk;
    const uint64_t u1 = std::min(n, u0 + kChunk);
    for (uint64_t u = u0; u < u1 && u + 1 < n; ++u) {
      uint64_t st = stream_seed(seed, u + 1);
      uint64_t v = u + 1;
      double p = std::min(w[u] * w[v] / S, 1.0);
      while (v < n && p > 0) {
        if (p != 1.0) {
          const double r = unit_open0(st);
          const double skip = std::floor(std::log(r) / std::log1p(-p));
          if (!(skip < static_cast<double>(n))) break;
          v += static_cast<uint64_t>(skip);
        }
        if (v < n) {
          const double q = std::min(w[u] * w[v] / S, 1.0);
          const double r = unit_open0(st);
          if (r < q / p) part.emplace_back(static_cast<uint32_t>(u), static_cast<uint32_t>(v));
          p = q;
          ++v;
        }
      }
    }
  }
  std::vector<std::pair<uint32_t, uint32_t>> edges;
  size_t total = 0;
  for (auto& part : parts) total += part.size();
  edges.reserve(total + n / 8);
  for (auto& part : parts) {
    edges.insert(edges.end(), part.begin(), part.end());
    std::vector<std::pair<uint32_t, uint32_t>>().swap(part);
  }
  connect_components(n, w, seed, edges);
  return build_csr(n, edges, out);
}

extern "C" int tgtp_gen_grid_rewired(uint32_t rows, uint32_t cols, double p_rewire,
                                     uint64_t seed, tgtp_csr* out) {
  if (rows < 2 || cols < 2 || !(p_rewire >= 0) || p_rewire > 1) return -1;
  const uint64_t n = static_cast<uint64_t>(rows) * cols;
  if (n > 0xFFFFFFFEull) return -1;
  std::vector<std::pair<uint32_t, uint32_t>> edges;
  edges.reserve(2 * n);
  for (uint32_t r = 0; r < rows; ++r)
    for (uint32_t c = 0; c + 1 < cols; ++c)
      edges.emplace_back(r * cols + c, r * cols + c + 1);
  for (uint32_t r = 0; r + 1 < rows; ++r)
    for (uint32_t c = 0; c < cols; ++c)
      edges.emplace_back(r * cols + c, (r + 1) * cols + c);
  const uint64_t E = edges.size();
  const uint64_t k = static_cast<uint64_t>(std::llround(p_rewire * static_cast<double>(E)));
  auto key = [](uint32_t a, uint32_t b) {
    if (a > b) std::swap(a, b);
    return (static_cast<uint64_t>(a) << 32) | b;
  };
  std::unordered_set<uint64_t> present;
  present.reserve(2 * E);
  for (auto& e : edges) present.insert(key(e.first, e.second));
  std::vector<uint64_t> order(E);
  std::iota(order.begin(), order.end(), 0ull);
  uint64_t st = stream_seed(seed, 0x6772696400ull);
  for (uint64_t i = 0; i < k; ++i) {  // partial Fisher-Yates: k distinct edges
    const uint64_t j = i + below(st, E - i);
    std::swap(order[i], order[j]);
  }
  for (uint64_t i = 0; i < k; ++i) {
    auto& e = edges[order[i]];
    const uint32_t a = e.first;
    present.erase(key(e.first, e.second));
    uint32_t c;
    do {
      c = static_cast<uint32_t>(below(st, n));
    } while (c == a || present.count(key(a, c)));
    e.second = c;
    present.insert(key(a, c));
  }
  std::vector<double> w(n, 1.0);
  connect_components(n, w, seed, edges);
  return build_csr(n, edges, out);
}

extern "C" int tgtp_write_graph_txt(const tgtp_csr* g, const char* path) {
  FILE* f = std::fopen(path, "w");
  if (!f) return -4;
  std::vector<char> buf(1 << 20);
  setvbuf(f, buf.data(), _IOFBF, buf.size());
  for (uint64_t u = 0; u < g->n; ++u) {
    bool loop_half = false;
    for (uint64_t p = g->idx[u]; p < g->idx[u + 1]; ++p) {
      const uint32_t v = g->adj[p];
      if (v < u) continue;
      if (v == u) {  // a self-loop is stored twice per input line
        loop_half = !loop_half;
        if (!loop_half) continue;
      }
      std::fprintf(f, "%llu %u\n", static_cast<unsigned long long>(u), v);
    }
  }
  return std::fclose(f) == 0 ? 0 : -4;
}

extern "C" int tgtp_csr_reverse_index(const tgtp_csr* g, uint32_t* rev) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < static_cast<int64_t>(g->n); ++u) {
    for (uint64_t p = g->idx[u]; p < g->idx[u + 1]; ++p) {
      const uint32_t v = g->adj[p];
      const uint32_t* first = std::lower_bound(g->adj + g->idx[v], g->adj + g->idx[v + 1],
                                               static_cast<uint32_t>(u));
      rev[p] = static_cast<uint32_t>(first - g->adj);
    }
  }
  return 0;
}
