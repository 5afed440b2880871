// graph_io.cpp -- host graph / eigenpair loading, seeded generators and CSR
// helpers of libtgtp_hip.so (declarations and semantics: tgtp_graph_io.h).
#include "tgtp_graph_io.h"

#include <fcntl.h>
#include <locale.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <type_traits>
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

// End of the characters that libstdc++'s `in >> x` gathers for a
// floating-point x (num_get::_M_extract_float) from p, blanks already
// skipped: [+-] digits [. digits] and, once a mantissa digit was seen,
// [eE] [+-] digits.  The value must be the whole of them (*whole: a mantissa
// digit and, after an exponent marker, an exponent digit), so "1e" and "1e+"
// fail where strtod would read 1, and "0x10" reads 0 and stops at 'x' where
// strtod would read a hexadecimal 16.
const char* float_token_end(const char* p, const char* end, bool* whole) {
  const char* q = p;
  if (q < end && (*q == '+' || *q == '-')) ++q;
  bool digits = false;
  while (q < end && *q >= '0' && *q <= '9') {
    ++q;
    digits = true;
  }
  if (q < end && *q == '.') {
    ++q;
    while (q < end && *q >= '0' && *q <= '9') {
      ++q;
      digits = true;
    }
  }
  *whole = digits;
  if (digits && q < end && (*q == 'e' || *q == 'E')) {
    ++q;
    if (q < end && (*q == '+' || *q == '-')) ++q;
    bool exp_digits = false;
    while (q < end && *q >= '0' && *q <= '9') {
      ++q;
      exp_digits = true;
    }
    *whole = exp_digits;
  }
  return q;
}

// The gathered field [p, q) converted as `in >> x` converts it in the
// classic locale: correctly rounded straight to T; an overflow (anything
// non-finite) fails, an underflow keeps strtod's value as libstdc++ does.
// *q is set to NUL for the conversion and restored, so [p, q] must be
// writable and no other thread may read *q meanwhile.
template <class T>
bool convert_float_field(char* p, char* q, T& out) {
  static const locale_t c_locale = newlocale(LC_ALL_MASK, "C", static_cast<locale_t>(0));
  if (!c_locale) return false;
  const char saved = *q;
  *q = '\0';
  char* next = nullptr;
  T v;
  if constexpr (std::is_same_v<T, float>)
    v = strtof_l(p, &next, c_locale);
  else
    v = strtod_l(p, &next, c_locale);
  *q = saved;
  if (next != q || !std::isfinite(v)) return false;
  out = v;
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

// One chunk of a text graph: its edges in file order, 1 + its largest id,
// and its first invalid line, if any (parsing stops there).
struct ParseChunk {
  std::vector<std::pair<uint32_t, uint32_t>> edges;
  uint64_t n = 0;
  const char* bad_line = nullptr;
  const char* bad_eol = nullptr;
};

// graph.h:29-97 line rules: '#' / '%' comment lines (also indented), blank
// lines skipped, two non-negative ids then anything.
void parse_lines(const char* p, const char* end, ParseChunk& c) {
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
      c.bad_line = line;
      c.bad_eol = eol;
      return;
    }
    c.edges.emplace_back(static_cast<uint32_t>(a), static_cast<uint32_t>(b));
    c.n = std::max<uint64_t>(c.n, static_cast<uint64_t>(std::max(a, b)) + 1);
  }
}

// Sorted-row CSR of both directions of every pair of every chunk: degree
// counts and row fills with atomics across threads, then each row sorted
// (the order of the fills does not matter after the sort).
int build_csr_chunks(uint64_t n, const std::vector<ParseChunk>& chunks, tgtp_csr* out) {
  uint64_t pairs = 0;
  for (const ParseChunk& c : chunks) pairs += c.edges.size();
  const uint64_t m = 2ull * pairs;
  if (m > 0xFFFFFFFFull) return -1;
  uint32_t* idx = static_cast<uint32_t*>(std::malloc((n + 1) * sizeof(uint32_t)));
  uint32_t* adj = static_cast<uint32_t*>(std::malloc((m ? m : 1) * sizeof(uint32_t)));
  if (!idx || !adj) {
    std::free(idx);
    std::free(adj);
    return -3;
  }
  std::vector<uint32_t> cnt(n, 0);
  const int nc = static_cast<int>(chunks.size());
  // (host threads are woken only when there is work for several: idle ones
  // spin for a while after a parallel loop and slow what follows)
#pragma omp parallel for schedule(static, 1) if (nc > 1)
  for (int t = 0; t < nc; ++t)
    for (const auto& e : chunks[static_cast<size_t>(t)].edges) {
      __atomic_fetch_add(&cnt[e.first], 1u, __ATOMIC_RELAXED);
      __atomic_fetch_add(&cnt[e.second], 1u, __ATOMIC_RELAXED);
    }
  idx[0] = 0;
  for (uint64_t v = 0; v < n; ++v) idx[v + 1] = idx[v] + cnt[v];
  std::copy(idx, idx + n, cnt.begin());  // next free slot per row
#pragma omp parallel for schedule(static, 1) if (nc > 1)
  for (int t = 0; t < nc; ++t)
    for (const auto& e : chunks[static_cast<size_t>(t)].edges) {
      adj[__atomic_fetch_add(&cnt[e.first], 1u, __ATOMIC_RELAXED)] = e.second;
      adj[__atomic_fetch_add(&cnt[e.second], 1u, __ATOMIC_RELAXED)] = e.first;
    }
#pragma omp parallel for schedule(dynamic, 4096) if (m >= (1ull << 20))
  for (int64_t v = 0; v < static_cast<int64_t>(n); ++v) std::sort(adj + idx[v], adj + idx[v + 1]);
  out->n = n;
  out->m = m;
  out->idx = idx;
  out->adj = adj;
  return 0;
}

// 64-bit digest of the binary cache's arrays: 8-byte little-endian words
// (the tail zero-padded), h = rotl(h ^ w, 29) * odd constant, finished with
// the byte count.
uint64_t digest_words(const void* p, size_t bytes, uint64_t h) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  size_t i = 0;
  for (; i + 8 <= bytes; i += 8) {
    uint64_t w;
    std::memcpy(&w, c + i, 8);
    h ^= w;
    h = ((h << 29) | (h >> 35)) * 0x9e3779b97f4a7c15ull;
  }
  if (i < bytes) {
    uint64_t w = 0;
    std::memcpy(&w, c + i, bytes - i);
    h ^= w;
    h = ((h << 29) | (h >> 35)) * 0x9e3779b97f4a7c15ull;
  }
  h ^= bytes;
  return h * 0xd6e8feb86659fd93ull;
}

constexpr char kCsrMagic[8] = {'T', 'G', 'T', 'P', 'C', 'S', 'R', '1'};
constexpr char kEigMagic[8] = {'T', 'G', 'T', 'P', 'E', 'I', 'G', '1'};

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
#pragma omp parallel for schedule(dynamic, 4096) if (m >= (1ull << 20))
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
  // The file is cut into chunks at line starts, parsed by all threads; a
  // line belongs to the chunk it starts in, so the first chunk holding an
  // invalid line holds the file's first invalid line.
  const char* data = f.data;
  const size_t size = f.size;
  const int threads = std::max(1, std::min<int>(omp_get_max_threads(),
                                                static_cast<int>(size >> 20) + 1));
  std::vector<const char*> cut(static_cast<size_t>(threads) + 1, data + size);
  cut[0] = data;
  for (int t = 1; t < threads; ++t) {
    const char* c = data + size / static_cast<size_t>(threads) * static_cast<size_t>(t);
    c = std::max(c, cut[static_cast<size_t>(t) - 1]);
    const char* nl = c < data + size
                         ? static_cast<const char*>(std::memchr(c, '\n', static_cast<size_t>(data + size - c)))
                         : nullptr;
    cut[static_cast<size_t>(t)] = nl ? nl + 1 : data + size;
  }
  std::vector<ParseChunk> chunks(static_cast<size_t>(threads));
#pragma omp parallel for schedule(static, 1) num_threads(threads)
  for (int t = 0; t < threads; ++t)
    parse_lines(cut[static_cast<size_t>(t)], cut[static_cast<size_t>(t) + 1], chunks[static_cast<size_t>(t)]);
  uint64_t n = 0, pairs = 0;
  for (const ParseChunk& c : chunks) {
    if (c.bad_line) {
      set_err(err, errlen, std::string("invalid edge in ") + path + ": " +
                               std::string(c.bad_line, static_cast<size_t>(c.bad_eol - c.bad_line)));
      return -1;
    }
    n = std::max(n, c.n);
    pairs += c.edges.size();
  }
  if (pairs == 0) {
    set_err(err, errlen, std::string("empty graph: ") + path);
    return -1;
  }
  if (2ull * pairs > 0xFFFFFFFFull) {
    set_err(err, errlen, "graph exceeds CSR index capacity");
    return -1;
  }
  const int rc = build_csr_chunks(n, chunks, out);
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

extern "C" int tgtp_save_csr_bin(const tgtp_csr* g, const char* path) {
  if (!g || !path || !g->idx || (g->m && !g->adj)) return -1;
  const std::string tmp = std::string(path) + ".tmp";
  FILE* fp = std::fopen(tmp.c_str(), "wb");
  if (!fp) return -4;
  const uint64_t hdr[2] = {g->n, g->m};
  const size_t ib = (g->n + 1) * sizeof(uint32_t), ab = g->m * sizeof(uint32_t);
  uint64_t h = digest_words(g->idx, ib, 0x6a09e667f3bcc908ull ^ g->n);
  h = digest_words(g->adj, ab, h ^ g->m);
  bool ok = std::fwrite(kCsrMagic, 1, 8, fp) == 8 && std::fwrite(hdr, 8, 2, fp) == 2 &&
            std::fwrite(g->idx, 1, ib, fp) == ib && (ab == 0 || std::fwrite(g->adj, 1, ab, fp) == ab) &&
            std::fwrite(&h, 8, 1, fp) == 1;
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return -4;
  }
  return 0;
}

extern "C" int tgtp_load_csr_bin(const char* path, tgtp_csr* out, char* err, size_t errlen) {
  if (!path || !out) {
    set_err(err, errlen, "null argument");
    return -1;
  }
  Mapped f;
  if (!f.open_file(path)) {
    set_err(err, errlen, std::string("cannot open CSR cache: ") + path);
    return -4;
  }
  const std::string bad = std::string("not a valid CSR cache: ") + path;
  uint64_t hdr[2];
  if (f.size < 24 + 4 + 8 || std::memcmp(f.data, kCsrMagic, 8) != 0) {
    set_err(err, errlen, bad);
    return -1;
  }
  std::memcpy(hdr, f.data + 8, 16);
  const uint64_t n = hdr[0], m = hdr[1];
  if (n >= 0xFFFFFFFFull || m > 0xFFFFFFFFull || f.size != 24 + (n + 1) * 4 + m * 4 + 8) {
    set_err(err, errlen, bad);
    return -1;
  }
  const char* pi = f.data + 24;
  const char* pa = pi + (n + 1) * 4;
  uint64_t h = digest_words(pi, (n + 1) * 4, 0x6a09e667f3bcc908ull ^ n);
  h = digest_words(pa, m * 4, h ^ m);
  uint64_t stored;
  std::memcpy(&stored, pa + m * 4, 8);
  if (h != stored) {
    set_err(err, errlen, bad + " (checksum)");
    return -1;
  }
  uint32_t* idx = static_cast<uint32_t*>(std::malloc((n + 1) * sizeof(uint32_t)));
  uint32_t* adj = static_cast<uint32_t*>(std::malloc((m ? m : 1) * sizeof(uint32_t)));
  if (!idx || !adj) {
    std::free(idx);
    std::free(adj);
    set_err(err, errlen, "out of host memory");
    return -3;
  }
  std::memcpy(idx, pi, (n + 1) * 4);
  if (m) std::memcpy(adj, pa, m * 4);
  // CSR invariants: rows in order, ending at m, ids in range, rows sorted
  bool ok = idx[0] == 0 && idx[n] == m;
  for (uint64_t v = 0; ok && v < n; ++v) ok = idx[v] <= idx[v + 1];
  int bad_rows = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : bad_rows) if (m >= (1ull << 20))
  for (int64_t v = 0; v < (ok ? static_cast<int64_t>(n) : 0); ++v)
    for (uint32_t p = idx[v]; p < idx[v + 1]; ++p)
      if (adj[p] >= n || (p > idx[v] && adj[p - 1] > adj[p])) {
        ++bad_rows;
        break;
      }
  if (!ok || bad_rows) {
    std::free(idx);
    std::free(adj);
    set_err(err, errlen, bad + " (CSR invariants)");
    return -1;
  }
  out->n = n;
  out->m = m;
  out->idx = idx;
  out->adj = adj;
  return 0;
}

extern "C" int tgtp_save_eigens_bin(const char* path, uint64_t n, int32_t omega,
                                    const double* values, const float* vectors) {
  if (!path || omega <= 0 || !values || (n && !vectors)) return -1;
  const uint64_t om = static_cast<uint64_t>(omega), cnt = n * om;
  for (uint64_t i = 0; i < om; ++i)
    if (!std::isfinite(values[i])) return -1;
  int nonfinite = 0;
#pragma omp parallel for schedule(static) reduction(+ : nonfinite) if (cnt >= (1ull << 20))
  for (int64_t i = 0; i < static_cast<int64_t>(cnt); ++i) nonfinite += !std::isfinite(vectors[i]);
  if (nonfinite) return -1;
  const std::string tmp = std::string(path) + ".tmp";
  FILE* fp = std::fopen(tmp.c_str(), "wb");
  if (!fp) return -4;
  const uint64_t hdr[2] = {n, om};
  const size_t vb = om * sizeof(double), xb = cnt * sizeof(float);
  uint64_t h = digest_words(values, vb, 0xbb67ae8584caa73bull ^ n);
  h = digest_words(vectors, xb, h ^ om);
  bool ok = std::fwrite(kEigMagic, 1, 8, fp) == 8 && std::fwrite(hdr, 8, 2, fp) == 2 &&
            std::fwrite(values, 1, vb, fp) == vb && (xb == 0 || std::fwrite(vectors, 1, xb, fp) == xb) &&
            std::fwrite(&h, 8, 1, fp) == 1;
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return -4;
  }
  return 0;
}

extern "C" int tgtp_load_eigens_bin(const char* path, uint64_t n, int32_t omega, double* values,
                                    float* vectors, char* err, size_t errlen) {
  if (!path || omega <= 0 || !values || (n && !vectors)) {
    set_err(err, errlen, "bad argument");
    return -1;
  }
  Mapped f;
  if (!f.open_file(path)) {
    set_err(err, errlen, std::string("cannot open eigenpair cache: ") + path);
    return -4;
  }
  const std::string bad = std::string("not a valid eigenpair cache: ") + path;
  const uint64_t om = static_cast<uint64_t>(omega);
  // sizes first, so that a foreign or truncated file is refused before any
  // product of its header is formed
  uint64_t hdr[2] = {0, 0};
  if (f.size >= 24) std::memcpy(hdr, f.data + 8, 16);
  if (f.size < 24 || std::memcmp(f.data, kEigMagic, 8) != 0 || hdr[0] != n || hdr[1] != om ||
      n > (UINT64_MAX / 4 - 64) / om || f.size != 24 + om * 8 + n * om * 4 + 8) {
    set_err(err, errlen, bad + " (expected n = " + std::to_string(n) +
                             ", omega = " + std::to_string(omega) + ")");
    return -1;
  }
  const char* pv = f.data + 24;
  const char* px = pv + om * 8;
  const size_t xb = n * om * 4;
  uint64_t h = digest_words(pv, om * 8, 0xbb67ae8584caa73bull ^ n);
  h = digest_words(px, xb, h ^ om);
  uint64_t stored;
  std::memcpy(&stored, px + xb, 8);
  if (h != stored) {
    set_err(err, errlen, bad + " (checksum)");
    return -1;
  }
  std::memcpy(values, pv, om * 8);
  if (xb) std::memcpy(vectors, px, xb);
  bool finite = true;
  for (uint64_t i = 0; i < om; ++i) finite = finite && std::isfinite(values[i]);
  int nonfinite = 0;
#pragma omp parallel for schedule(static) reduction(+ : nonfinite) if (n * om >= (1ull << 20))
  for (int64_t i = 0; i < static_cast<int64_t>(n * om); ++i) nonfinite += !std::isfinite(vectors[i]);
  if (!finite || nonfinite) {
    set_err(err, errlen, bad + " (non-finite values)");
    return -1;
  }
  return 0;
}

extern "C" int tgtp_load_eigens_txt(const char* path, uint64_t n, int32_t omega,
                                    double* values, float* vectors, char* err,
                                    size_t errlen) {
  Mapped f;
  if (!f.open_file(path)) {
    set_err(err, errlen, std::string("cannot open eigenpairs: ") + path);
    return -4;
  }
  if (omega <= 0) return 0;
  // the conversions need a terminated, writable buffer; copy once
  std::string buf(f.data ? f.data : "", f.size);
  char* const data = buf.data();
  const size_t size = buf.size();
  // The reference reads omega (n + 1) blank-separated fields in order, each
  // eigenvalue as double and each entry as float, and stops at the first
  // that does not read.  Here the file is cut just after blanks (no field
  // spans a cut, and the sequential read is between fields at each) into
  // chunks that all host threads parse: pass 1 counts each chunk's fields
  // up to the first that cannot convert, pass 2 converts each field to the
  // type its index in the whole file calls for.  The lowest failing index
  // names the error the sequential read would have met first.
  const uint64_t per = n + 1, need = per * static_cast<uint64_t>(omega);
  const int threads = std::max(1, std::min<int>(omp_get_max_threads(),
                                                static_cast<int>(size >> 18) + 1));
  std::vector<char*> cut(static_cast<size_t>(threads) + 1, data + size);
  cut[0] = data;
  for (int t = 1; t < threads; ++t) {
    char* c = std::max(data + size / static_cast<size_t>(threads) * static_cast<size_t>(t),
                       cut[static_cast<size_t>(t) - 1]);
    while (c < data + size && !is_space(*c)) ++c;
    cut[static_cast<size_t>(t)] = c < data + size ? c + 1 : data + size;
  }
  std::vector<uint64_t> count(static_cast<size_t>(threads), 0), bad(static_cast<size_t>(threads), UINT64_MAX);
  std::vector<uint8_t> stopped(static_cast<size_t>(threads), 0);
#pragma omp parallel for schedule(static, 1) num_threads(threads)
  for (int t = 0; t < threads; ++t) {
    const char* p = cut[static_cast<size_t>(t)];
    const char* const e = cut[static_cast<size_t>(t) + 1];
    uint64_t c = 0;
    for (;;) {
      while (p < e && is_space(*p)) ++p;
      if (p == e) break;
      bool whole = false;
      const char* q = float_token_end(p, e, &whole);
      if (!whole) {
        stopped[static_cast<size_t>(t)] = 1;
        break;
      }
      ++c;
      p = q;
    }
    count[static_cast<size_t>(t)] = c;
  }
  // index of the first field that cannot be read at all (malformed, or past
  // the end of the file); chunks after a stopped one do not count
  std::vector<uint64_t> first(static_cast<size_t>(threads), 0);
  uint64_t unreadable = 0;
  bool hit = false;
  for (int t = 0; t < threads; ++t) {
    first[static_cast<size_t>(t)] = unreadable;
    if (hit) continue;
    unreadable += count[static_cast<size_t>(t)];
    hit = stopped[static_cast<size_t>(t)] != 0;
  }
  const uint64_t lim = std::min(need, unreadable);
#pragma omp parallel for schedule(static, 1) num_threads(threads)
  for (int t = 0; t < threads; ++t) {
    uint64_t idx = first[static_cast<size_t>(t)];
    char* p = cut[static_cast<size_t>(t)];
    for (uint64_t k = 0; k < count[static_cast<size_t>(t)] && idx < lim; ++k, ++idx) {
      while (is_space(*p)) ++p;
      bool whole = false;
      char* q = const_cast<char*>(float_token_end(p, cut[static_cast<size_t>(t) + 1], &whole));
      const uint64_t i = idx / per, r = idx % per;
      bool ok;
      if (r == 0) {
        ok = convert_float_field(p, q, values[i]);
      } else {
        // parsed straight to float, as aesc.h:60
        ok = convert_float_field(p, q, vectors[static_cast<uint64_t>(omega) * (r - 1) + i]);
      }
      if (!ok) {
        bad[static_cast<size_t>(t)] = idx;
        break;
      }
      p = q;
    }
  }
  uint64_t first_bad = unreadable;
  for (int t = 0; t < threads; ++t) first_bad = std::min(first_bad, bad[static_cast<size_t>(t)]);
  if (first_bad < need) {
    set_err(err, errlen, std::string(first_bad % per == 0 ? "invalid eigenvalue in "
                                                           : "invalid eigenvector in ") +
                             path);
    return -1;
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
#pragma omp parallel for schedule(dynamic, 1024) if (g->m >= (1ull << 20))
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
