/*
 * tgtp_graph_io.h -- host-side graph and eigenpair I/O plus the seeded
 * synthetic graph generators the benchmarks run on.  Part of libtgtp_hip.so;
 * no GPU involvement.
 *
 * Reference interfaces replaced:
 *   tgtp_load_txt         <- fuser::graph_t::load_txt(filename)   src/graph.h:29-97
 *   tgtp_load_eigens_txt  <- EigenVecs<float>::load_txt(path)     src/aesc.h:51-64
 *   tgtp_check_aesc_graph <- AESCGraph ctor validation            src/aesc.h:85-92
 * The generators have no reference counterpart: the reference downloads
 * SNAP graphs (datasets/urls.txt), which needs a network.
 */
#ifndef TGTP_GRAPH_IO_H
#define TGTP_GRAPH_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Owned CSR (free with tgtp_csr_free). */
typedef struct tgtp_csr {
  uint64_t n, m;
  uint32_t* idx; /* [n+1] */
  uint32_t* adj; /* [m]   */
} tgtp_csr;

/* graph.h:29-97 semantics: '#'/'%' comment lines, two non-negative ids per
 * line (extra fields ignored), both directions stored, duplicates kept,
 * self-loops stored twice, n = max id + 1, rows sorted.  Errors: -1 invalid
 * edge / empty graph / CSR overflow, -4 cannot open.  err (may be NULL)
 * receives a message of at most errlen bytes. */
int tgtp_load_txt(const char* path, tgtp_csr* out, char* err, size_t errlen);
void tgtp_csr_free(tgtp_csr* g);

/* Binary CSR cache -- no reference counterpart (the reference parses the
 * text on every run, graph.h:29-97); for callers that load the same large
 * graph repeatedly.  Layout, little-endian: "TGTPCSR1", u64 n, u64 m,
 * u32 idx[n+1], u32 adj[m], u64 digest of the two arrays.  Save writes a
 * temporary file and renames it: 0, -1 bad argument, -4 write failure.
 * Load checks the size, the digest and the CSR invariants (idx from 0 to m
 * non-decreasing, ids < n, rows sorted): 0, -1 damaged or foreign file,
 * -3 out of host memory, -4 cannot open. */
int tgtp_save_csr_bin(const tgtp_csr* g, const char* path);
int tgtp_load_csr_bin(const char* path, tgtp_csr* out, char* err, size_t errlen);

/* aesc.h:51-64: omega lines "lambda x_0 ... x_{n-1}", read as omega (n+1)
 * blank-separated fields exactly as `in >> x` reads them in the classic
 * locale (eigenvalues as double, entries directly as float; a field that
 * does not read whole, or overflows, is an error -1, as is a missing field;
 * missing file -4); parsed by all host threads.  values [omega] double,
 * vectors [n*omega] float vertex-major. */
int tgtp_load_eigens_txt(const char* path, uint64_t n, int32_t omega,
                         double* values, float* vectors, char* err,
                         size_t errlen);

/* Binary eigenpair cache -- no reference counterpart (the reference parses
 * sorted_eigens_<omega>.txt on every run, aesc.h:51-64: about 800 MB of text
 * at n = 262,144, omega = 128).  Layout, little-endian: "TGTPEIG1", u64 n,
 * u64 omega, f64 values[omega], f32 vectors[n*omega] vertex-major, u64
 * digest of the two arrays -- the arrays tgtp_load_eigens_txt returns.
 * Save refuses non-finite values (-1) and writes a temporary file then
 * renames it (-4 write failure).  Load takes the n and omega the caller
 * expects: a file of other sizes, damaged, or holding non-finite values is
 * -1; -4 cannot open. */
int tgtp_save_eigens_bin(const char* path, uint64_t n, int32_t omega,
                         const double* values, const float* vectors);
int tgtp_load_eigens_bin(const char* path, uint64_t n, int32_t omega,
                         double* values, float* vectors, char* err,
                         size_t errlen);

/* aesc.h:87-89: 2 <= omega <= n and no vertex of degree 0.  0 or -1. */
int tgtp_check_aesc_graph(const tgtp_csr* g, int32_t omega, char* err,
                          size_t errlen);

/* Seeded Chung-Lu graph, undirected, simple (no self-loops, no duplicates):
 * expected degrees w_i = c*(i+1)^(-1/(beta-1)) scaled to mean `mean_deg`,
 * capped at w_cap when w_cap > 0, edge (i,j) with probability
 * min(1, w_i*w_j/sum(w)), sampled in O(n+m) by geometric skipping
 * (Miller & Hagberg, "Efficient generation of networks with given expected
 * degrees", WAW 2011).  Every vertex outside the largest connected component
 * is then joined to it by one edge from its smallest-id vertex to a
 * weight-sampled vertex of that component, so the result is connected, has
 * no isolated vertex and keeps n exactly.  Deterministic in (n, beta,
 * mean_deg, w_cap, seed) and independent of thread count. */
int tgtp_gen_chung_lu(uint64_t n, double beta, double mean_deg, double w_cap,
                      uint64_t seed, tgtp_csr* out);

/* rows x cols 4-neighbour grid; round(p_rewire * E) edges chosen by seed
 * keep their first endpoint and move the other to a uniform vertex,
 * redrawing self-loops and duplicates; connectivity repaired as above.
 * E = rows*(cols-1) + (rows-1)*cols is preserved. */
int tgtp_gen_grid_rewired(uint32_t rows, uint32_t cols, double p_rewire,
                          uint64_t seed, tgtp_csr* out);

/* Writes the undirected edge list "u v" (u <= v, CSR order, one line per
 * CSR entry with u <= v) -- the datasets/README.md:3-12 graph.txt format. */
int tgtp_write_graph_txt(const tgtp_csr* g, const char* path);

/* rev[pos(u,v)] = position of the first u in v's row (the lower_bound of
 * aesc.h:207-209 and 541-543).  rev [m]. */
int tgtp_csr_reverse_index(const tgtp_csr* g, uint32_t* rev);

#ifdef __cplusplus
}
#endif
#endif /* TGTP_GRAPH_IO_H */
