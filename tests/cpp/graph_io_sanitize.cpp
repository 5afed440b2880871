// Host loaders of libtgtp_hip.so (tgtp_amd/csrc/graph_io.cpp), built here
// with AddressSanitizer and UBSan, over crafted and seeded-random inputs:
//   * text graphs: tgtp_load_txt against the oracle's istream loader
//     (orc_load_txt, graph.h:29-97): same success, same CSR, or both fail;
//     many-thread parses of a large file with the same inputs;
//   * eigenpair files: tgtp_load_eigens_txt against orc_load_eigens
//     (aesc.h:51-64): same success and bitwise the same values, on small
//     files and on files large enough for a many-thread parse;
//   * binary CSR and eigenpair caches: every truncation and seeded byte
//     damage of a saved cache is refused.
// usage: graph_io_sanitize <dir> <cases>; prints "ok <cases> ..." (with the
// number of inputs each loader accepted or refused) on success.
#include <omp.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../include/tgtp_graph_io.h"
#include "../../oracle/oracle.h"

namespace {

int g_fail = 0;
int g_graphs_ok = 0, g_eigens_ok = 0, g_caches_refused = 0, g_eig_caches_refused = 0;

void fail(const std::string& what, const std::string& input) {
  if (g_fail++ < 10) std::printf("MISMATCH %s\n--- input ---\n%s\n---\n", what.c_str(), input.c_str());
}

void write_file(const std::string& path, const std::string& s) {
  std::ofstream o(path, std::ios::binary);
  o.write(s.data(), static_cast<std::streamsize>(s.size()));
}

// Pieces of a graph line: small ids, signs, separators, comments and the
// malformed forms a loader must refuse (or, for trailing fields, ignore).
const char* const kGraphTok[] = {
    "0", "1", "2", "3", "7", "12", "31", "+4", "-0", "-1", "007", "0x10", "3x", "1.5", "1e2",
    "4294967295", "99999999999999999999", "-9223372036854775808", "#", "%", "a", "\t", " ",
    "  ", "\r", "\v", "\f", ""};

std::string random_graph_text(std::mt19937_64& rng) {
  std::string s;
  const int lines = static_cast<int>(rng() % 12);
  const int ntok = static_cast<int>(sizeof(kGraphTok) / sizeof(kGraphTok[0]));
  for (int l = 0; l < lines; ++l) {
    const uint64_t kind = rng() % 10;
    if (kind < 6) {  // a well-formed edge, maybe with extra fields
      s += std::to_string(rng() % 40) + (rng() % 2 ? " " : "\t") + std::to_string(rng() % 40);
      if (rng() % 4 == 0) s += std::string(" ") + kGraphTok[rng() % ntok];
    } else {  // random tokens
      const int k = static_cast<int>(rng() % 4);
      for (int i = 0; i < k; ++i) {
        s += kGraphTok[rng() % ntok];
        if (rng() % 3) s += ' ';
      }
    }
    if (l + 1 < lines || rng() % 2) s += rng() % 8 == 0 ? "\r\n" : "\n";
  }
  return s;
}

bool check_csr(const tgtp_csr& g) {
  if (g.idx[0] != 0 || g.idx[g.n] != g.m) return false;
  for (uint64_t v = 0; v < g.n; ++v) {
    if (g.idx[v + 1] < g.idx[v]) return false;
    for (uint64_t p = g.idx[v]; p < g.idx[v + 1]; ++p) {
      if (g.adj[p] >= g.n) return false;
      if (p > g.idx[v] && g.adj[p] < g.adj[p - 1]) return false;
    }
  }
  return true;
}

void compare_graph(const std::string& path, const std::string& text) {
  tgtp_csr g{};
  char err[256];
  const int rc = tgtp_load_txt(path.c_str(), &g, err, sizeof(err));
  uint64_t n = 0, m = 0;
  uint32_t *idx = nullptr, *adj = nullptr;
  const int orc = orc_load_txt(path.c_str(), &n, &m, &idx, &adj);
  if ((rc == 0) != (orc == 0)) {
    fail("graph: product rc " + std::to_string(rc) + ", oracle rc " + std::to_string(orc), text);
  } else if (rc == 0) {
    ++g_graphs_ok;
    if (!check_csr(g)) fail("graph: CSR invariants", text);
    if (g.n != n || g.m != m ||
        std::memcmp(g.idx, idx, (n + 1) * sizeof(uint32_t)) != 0 ||
        (m && std::memcmp(g.adj, adj, m * sizeof(uint32_t)) != 0))
      fail("graph: CSR differs", text);
  }
  if (rc == 0) tgtp_csr_free(&g);
  if (orc == 0) {
    orc_free(idx);
    orc_free(adj);
  }
}

const char* const kEigTok[] = {
    "0", "1", "-1", "0.5", "-.25", ".5", "5.", "1e-3", "2.5E+2", "1e-40", "1e-400", "1e40",
    "3.4028236e38", "1e", "1e+", "0x10", "0x1p3", "nan", "inf", "-inf", "1.2.3", "1e5e5",
    "+", ".", "-0", "00012", "abc", "\n", " ", "\t", "\r\n"};

std::string random_eigen_text(std::mt19937_64& rng, int values) {
  std::string s;
  const int ntok = static_cast<int>(sizeof(kEigTok) / sizeof(kEigTok[0]));
  for (int i = 0; i < values; ++i) {
    if (rng() % 6 == 0) {
      s += kEigTok[rng() % ntok];
    } else {
      char b[64];
      std::snprintf(b, sizeof(b), "%.9g", std::uniform_real_distribution<double>(-2, 2)(rng));
      s += b;
    }
    s += rng() % 5 == 0 ? "\n" : " ";
  }
  if (rng() % 3 == 0) s.resize(rng() % (s.size() + 1));
  return s;
}

void compare_eigens(const std::string& path, const std::string& text, uint64_t n, int omega) {
  std::vector<double> l1(omega, -7.0), l2(omega, -7.0);
  std::vector<float> x1(n * omega, -7.0f), x2(n * omega, -7.0f);
  char err[256];
  const int rc = tgtp_load_eigens_txt(path.c_str(), n, omega, l1.data(), x1.data(), err, sizeof(err));
  const int orc = orc_load_eigens(path.c_str(), n, omega, l2.data(), x2.data());
  if ((rc == 0) != (orc == 0)) {
    fail("eigens: product rc " + std::to_string(rc) + ", oracle rc " + std::to_string(orc), text);
  } else if (rc == 0 && (std::memcmp(l1.data(), l2.data(), l1.size() * sizeof(double)) != 0 ||
                         std::memcmp(x1.data(), x2.data(), x1.size() * sizeof(float)) != 0)) {
    fail("eigens: values differ", text);
  } else if (rc == 0) {
    ++g_eigens_ok;
  }
}

// A file of omega (n + 1) fields large enough for a many-thread parse, with
// up to two odd tokens at random places (inside the fields read, or past
// them) and sometimes cut short.
std::string large_eigen_text(std::mt19937_64& rng, uint64_t n, int omega) {
  const int ntok = static_cast<int>(sizeof(kEigTok) / sizeof(kEigTok[0]));
  std::vector<std::string> fields;
  const uint64_t total = (n + 1) * static_cast<uint64_t>(omega) + rng() % 3;
  char b[64];
  for (uint64_t i = 0; i < total; ++i) {
    std::snprintf(b, sizeof(b), "%.9g", std::uniform_real_distribution<double>(-1, 1)(rng));
    fields.emplace_back(b);
  }
  const int odd = static_cast<int>(rng() % 3);
  for (int k = 0; k < odd; ++k) fields[rng() % fields.size()] = kEigTok[rng() % ntok];
  std::string s;
  for (uint64_t i = 0; i < fields.size(); ++i) {
    s += fields[i];
    s += (i + 1) % (n + 1) == 0 ? "\n" : (rng() % 50 == 0 ? "\t " : " ");
  }
  if (rng() % 4 == 0) s.resize(s.size() - rng() % 64);
  return s;
}

void damaged_caches(const std::string& dir, std::mt19937_64& rng, int cases) {
  tgtp_csr g{};
  if (tgtp_gen_chung_lu(300, 2.5, 6.0, 0.0, 11, &g) != 0) {
    fail("cache: generator", "");
    return;
  }
  const std::string path = dir + "/g.csr";
  if (tgtp_save_csr_bin(&g, path.c_str()) != 0) {
    fail("cache: save", "");
    return;
  }
  std::ifstream in(path, std::ios::binary);
  const std::string good((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  tgtp_csr h{};
  char err[256];
  if (tgtp_load_csr_bin(path.c_str(), &h, err, sizeof(err)) != 0 || h.n != g.n || h.m != g.m ||
      std::memcmp(h.adj, g.adj, g.m * sizeof(uint32_t)) != 0)
    fail("cache: round trip", "");
  else
    tgtp_csr_free(&h);
  const std::string bad = dir + "/bad.csr";
  for (int c = 0; c < cases; ++c) {
    std::string s = good;
    if (c % 3 == 0) {
      s.resize(rng() % s.size());
    } else {
      const int flips = 1 + static_cast<int>(rng() % 3);
      for (int f = 0; f < flips; ++f) s[rng() % s.size()] ^= static_cast<char>(1 + rng() % 255);
    }
    if (s == good) continue;
    write_file(bad, s);
    tgtp_csr b{};
    if (tgtp_load_csr_bin(bad.c_str(), &b, err, sizeof(err)) == 0) {
      fail("cache: damaged file accepted (case " + std::to_string(c) + ")", "");
      tgtp_csr_free(&b);
    } else {
      ++g_caches_refused;
    }
  }
  tgtp_csr_free(&g);
}

void damaged_eigen_caches(const std::string& dir, std::mt19937_64& rng, int cases) {
  const uint64_t n = 300;
  const int omega = 8;
  std::vector<double> lam(omega);
  std::vector<float> X(n * omega);
  for (auto& v : lam) v = std::uniform_real_distribution<double>(-1, 1)(rng);
  for (auto& v : X) v = std::uniform_real_distribution<float>(-1, 1)(rng);
  const std::string path = dir + "/e.bin";
  if (tgtp_save_eigens_bin(path.c_str(), n, omega, lam.data(), X.data()) != 0) {
    fail("eigen cache: save", "");
    return;
  }
  std::ifstream in(path, std::ios::binary);
  const std::string good((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  std::vector<double> l2(omega);
  std::vector<float> X2(n * omega);
  char err[256];
  if (tgtp_load_eigens_bin(path.c_str(), n, omega, l2.data(), X2.data(), err, sizeof(err)) != 0 ||
      l2 != lam || X2 != X)
    fail("eigen cache: round trip", "");
  if (tgtp_load_eigens_bin(path.c_str(), n + 1, omega, l2.data(), X2.data(), err, sizeof(err)) == 0)
    fail("eigen cache: accepted for another n", "");
  const std::string bad = dir + "/bad.bin";
  for (int c = 0; c < cases; ++c) {
    std::string s = good;
    if (c % 3 == 0) {
      s.resize(rng() % s.size());
    } else {
      const int flips = 1 + static_cast<int>(rng() % 3);
      for (int f = 0; f < flips; ++f) s[rng() % s.size()] ^= static_cast<char>(1 + rng() % 255);
    }
    if (s == good) continue;
    write_file(bad, s);
    if (tgtp_load_eigens_bin(bad.c_str(), n, omega, l2.data(), X2.data(), err, sizeof(err)) == 0)
      fail("eigen cache: damaged file accepted (case " + std::to_string(c) + ")", "");
    else
      ++g_eig_caches_refused;
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const std::string dir = argv[1];
  const int cases = std::atoi(argv[2]);
  std::mt19937_64 rng(20240611);
  const std::string path = dir + "/t.txt";
  for (int c = 0; c < cases; ++c) {
    const std::string text = random_graph_text(rng);
    write_file(path, text);
    compare_graph(path, text);
  }
  // a file large enough for a many-thread parse, with an invalid line late
  {
    std::string big;
    for (int i = 0; i < 300000; ++i)
      big += std::to_string(rng() % 5000) + " " + std::to_string(rng() % 5000) + "\n";
    write_file(path, big);
    compare_graph(path, "<large file>");
    big += "17 x\n1 2\n";
    write_file(path, big);
    compare_graph(path, "<large file, invalid line at the end>");
  }
  for (int c = 0; c < cases; ++c) {
    const uint64_t n = 1 + rng() % 5;
    const int omega = 1 + static_cast<int>(rng() % 3);
    const std::string text = random_eigen_text(rng, static_cast<int>(omega * (n + 1)));
    write_file(path, text);
    compare_eigens(path, text, n, omega);
  }
  for (int c = 0; c < 16 + cases / 250; ++c) {
    const uint64_t n = 20000 + rng() % 40000;
    const int omega = 2 + static_cast<int>(rng() % 3);
    const std::string text = large_eigen_text(rng, n, omega);
    write_file(path, text);
    compare_eigens(path, "<large eigenpair file " + std::to_string(c) + ">", n, omega);
  }
  damaged_caches(dir, rng, cases);
  damaged_eigen_caches(dir, rng, cases);
  if (g_fail) {
    std::printf("%d mismatches\n", g_fail);
    return 1;
  }
  std::printf(
      "ok %d graphs_parsed %d eigen_files_parsed %d damaged_caches_refused %d "
      "damaged_eigen_caches_refused %d\n",
      cases, g_graphs_ok, g_eigens_ok, g_caches_refused, g_eig_caches_refused);
  return 0;
}
