/*
 * oracle.h -- CPU restatement of the TGT+/SABA hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (tgtp_amd/, include/)
 * may include, link or call this.  It is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
 * checker, never as the thing measured or shipped.
 *
 * Each function cites the reference lines it restates (paths relative to the
 * reference repository root).  Floating-point expressions keep the
 * reference's exact mix of float and double so that a build with
 * -ffp-contract=off -fno-fast-math reproduces an IEEE build of the reference
 * bit for bit.
 *
 * Pinning (how each function is checked against something other than itself):
 *   orc_saba_walks_avx2: equal to orc_saba_walks on every test graph,
 *       including the fallback-forcing known answers and a non-symmetric CSR.
 *   orc_fmix32 / orc_splitmix32 / orc_saba_walks:
 *       recorded known answers of the reference's own saba_randomwalk
 *       (C9 at 40 hops and K5 at 20 hops, both forcing the whole-bouquet
 *       SplitMix32 fallback) plus closed-form invariants.
 *   orc_load_txt / orc_load_eigens / orc_taus / orc_tgtp (walk mode MT19937):
 *       bit-exact against the reference's own aesc_baseline.h and graph.h,
 *       compiled from the reference tree by oracle/Makefile (target ref).
 *   orc_mt_walks variant 0 (regular): per-vertex counts equal the reference's
 *       own scalar_randomwalk (rw_experiment.cpp, compiled by oracle/Makefile,
 *       target ref); the generator is std::mt19937, checked by the standard's
 *       known answer (orc_mt19937_nth(5489, 10000) == 4123659995).
 *   orc_mt_walks variant 1 (simd): the reference's simd_randomwalk needs
 *       Intel SVML (_mm256_rem_epu32), so it is not built; at depth 1 both
 *       variants draw the same outputs and must agree (tested); at depth > 1
 *       the draw order (8 draws per hop in lane order,
 *       rw_experiment_simd.cpp:55-57) is pinned by brute-force restatements
 *       on std::mt19937 streams restated in numpy, on graphs where the
 *       regular variant's order gives other counts.
 *   orc_saba_walks / orc_mt_walks on a CSR without both directions of every
 *       edge (the reference divides by zero there): a walker at a vertex
 *       without neighbours stops; checked against brute-force Python
 *       restatements of that rule (tests/test_oracle_pinning.py,
 *       tests/test_mt_walks.py).
 *   orc_tgtp (walk mode SABA): everything except the 2-way walk is the code
 *       path pinned above.  Phase 1 of the 2-way walk is checked against the
 *       pinned 1-way walk (same grid, hash and multiply-high step).
 *       The fallback phase of the 2-way walk (two successive SplitMix32 draws
 *       from one state, aesc.h:339-340, the first for the walker from s):
 *       pinned by a brute-force restatement of aesc.h:241-391 on graphs of
 *       power-of-two degrees (a 16-leaf star, the 4-cube), where every
 *       bouquet falls back and the restatement with the draws swapped gives
 *       a different sum (the oracle with its draws swapped fails that test).
 */
#ifndef TGTP_ORACLE_H
#define TGTP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* MurmurHash3 fmix32 (aesc.h:31-41; rw_experiment_saba.cpp:44-51). */
uint32_t orc_fmix32(uint32_t h);
/* SplitMix32 16/15/15 mixer; advances *state (aesc.h:23-30). */
uint32_t orc_splitmix32(uint32_t* state);

/* SABA 1-way walks (rw_experiment_saba.cpp:65-145) for every source in
 * [src_begin, src_end) with nonzero degree; counters[v] += visits (uint32,
 * wrapping).  OpenMP schedule(dynamic, chunk) over sources as in
 * rw_experiment_saba.cpp:194; threads <= 0 keeps the OpenMP default.
 * Returns 0, or -1 for n_walks not a multiple of 8 / above 2^32-1. */
int orc_saba_walks(uint64_t n, const uint32_t* idx, const uint32_t* adj,
                   uint32_t src_begin, uint32_t src_end, uint32_t depth,
                   uint64_t n_walks, uint32_t* counters, int threads,
                   int chunk);

/* The same walks with the 8 walkers of a bouquet in one AVX2 register
 * (orc_walk_avx2.cpp): equal counters, the host CPU baseline of the walk
 * engine.  Returns -2 where the CPU lacks AVX2, -1 for n_walks not a
 * multiple of 8 or a graph beyond 2^31 vertices / entries. */
int orc_saba_walks_avx2(uint64_t n, const uint32_t* idx, const uint32_t* adj,
                        uint32_t src_begin, uint32_t src_end, uint32_t depth,
                        uint64_t n_walks, uint32_t* counters, int threads, int chunk);

/* One source's walks into a caller counter array (single thread). */
void orc_saba_walk_source(const uint32_t* idx, const uint32_t* adj,
                          uint32_t src, uint32_t depth, uint64_t n_walks,
                          uint32_t* counters);

/* MT19937 comparator walks: variant 0 rw_experiment.cpp:34-47, variant 1
 * rw_experiment_simd.cpp:44-82; per-source std::mt19937 seeded 42 + src
 * (rw_experiment.cpp:99-104); counters[v] += visits.  -1 for an unknown
 * variant or (variant 1) n_walks not a multiple of 8. */
int orc_mt_walks(uint64_t n, const uint32_t* idx, const uint32_t* adj,
                 uint32_t src_begin, uint32_t src_end, uint32_t depth,
                 uint64_t n_walks, int variant, uint32_t* counters, int threads);
/* k-th output (1-based) of std::mt19937(seed). */
uint32_t orc_mt19937_nth(uint32_t seed, uint64_t k);

/* SABA 2-way walk of aesc.h:241-391; returns the lane-summed float. */
float orc_walk2_saba(const uint32_t* idx, const uint32_t* adj, uint32_t s,
                     uint32_t v, uint64_t nr, uint32_t len,
                     const float* pvec);

/* tau bounds (aesc.h:129-216), eps_half = float(epsilon / 2). */
int orc_taus(uint64_t n, uint64_t m, const uint32_t* idx, const uint32_t* adj,
             int omega, const double* lam, const float* X, float eps_half,
             int32_t* taus, int threads);

enum { ORC_WALK_SABA = 0, ORC_WALK_MT19937 = 1 };

/* Optional per-stage outputs of orc_tgtp (any pointer may be NULL). */
typedef struct {
  int32_t* taus;        /* [m] */
  uint64_t* nr;         /* [m] walk count per CSR position, 0 if no walks */
  float* walk_sum;      /* [m] raw 2-way walk sum x per position */
  float* preds_row;     /* [m] preds before symmetrization */
  int32_t* ell;         /* [n] rounds of diffusion per source */
  float* edge_max;      /* [n] */
  float* gmax;          /* [n] */
  float* gmin;          /* [n] */
  uint32_t* q_size;     /* [n] final frontier size */
} orc_tgtp_trace;

/* Full TGT+ (AESC::tgtp, aesc.h:120-127 and 444-551).  walk_mode selects
 * the SABA walk of aesc.h or the MT19937 walk of aesc_baseline.h:363-379
 * (the latter forces one thread, sources in order, one global generator
 * seeded 42, as aesc_baseline.h:17,261 do).  scores: [m]. */
int orc_tgtp(uint64_t n, uint64_t m, const uint32_t* idx, const uint32_t* adj,
             int omega, const double* lam, const float* X, double epsilon,
             double delta, int gamma, int walk_mode, int threads,
             float* scores, orc_tgtp_trace* trace);

/* Same per-source pipeline restricted to sources [src_begin, src_end) and
 * no symmetrization: used to time a bounded sample on the host.  Writes
 * only the rows of those sources in preds_rows[m]. */
int orc_tgtp_rows(uint64_t n, uint64_t m, const uint32_t* idx,
                  const uint32_t* adj, const int32_t* taus, double epsilon,
                  double delta, int gamma, uint32_t src_begin,
                  uint32_t src_end, int threads, float* preds_rows);

/* Same for an arbitrary list of sources (count entries, each < n); the
 * score of an edge whose endpoints are both listed is
 * preds_rows[pos] + preds_rows[first reverse position] (aesc.h:532-550). */
int orc_tgtp_sources(uint64_t n, uint64_t m, const uint32_t* idx,
                     const uint32_t* adj, const int32_t* taus, double epsilon,
                     double delta, int gamma, const uint32_t* sources,
                     uint64_t count, int threads, float* preds_rows);

/* graph.h:29-97 (undirected, skip = 0).  Allocates *idx [n+1] and *adj [m]
 * with malloc; caller frees with orc_free.  Returns 0 or -1 (message in
 * orc_last_error). */
int orc_load_txt(const char* path, uint64_t* n, uint64_t* m, uint32_t** idx,
                 uint32_t** adj);
/* aesc.h:51-64: lam [omega] double, X [n*omega] float vertex-major. */
int orc_load_eigens(const char* path, uint64_t n, int omega, double* lam,
                    float* X);
void orc_free(void* p);
const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
