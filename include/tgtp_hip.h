/*
 * tgtp_hip.h -- C ABI of the MI355X (gfx950) TGT+/SABA hot path.
 *
 * Plain C: fixed-width integers, plain pointers and sizes, int status codes,
 * no exceptions across the boundary, caller-owned outputs.  One context per
 * host thread; calls on one context are serialized by the caller.
 *
 * Reference interfaces replaced (ggokturk/tgtp):
 *   tgtp_score          <- float* AESC::tgtp()            src/aesc.h:120-127
 *                          (AESC(g, config) ctor           src/aesc.h:113-119)
 *   tgtp_saba_walks     <- saba_randomwalk(src, len, n_walk, pvec, g)
 *                          + its driver loop and reduction src/rw_experiment_saba.cpp:65-145, 190-207
 *   tgtp_taus           <- AESC::calculate_taus(eps)      src/aesc.h:164-216
 *   tgtp_graph_*        <- graph_t CSR (idx, adj, n, m)   src/graph.h:10-28
 *   tgtp_load_txt       <- graph_t::load_txt               src/graph.h:29-97   (tgtp_graph_io.h)
 *   tgtp_load_eigens_txt<- EigenVecs<float>::load_txt      src/aesc.h:51-64    (tgtp_graph_io.h)
 *
 * Data layout (host and device): idx uint32[n+1], adj uint32[m], rows sorted
 * ascending, both directions of every undirected edge, duplicates kept
 * (graph.h:13-15, 63-66).  Eigenvectors float[n*omega] vertex-major
 * (X[v*omega + i] = component v of eigenvector i, aesc.h:60), eigenvalues
 * double[omega] sorted by descending magnitude.
 */
#ifndef TGTP_HIP_H
#define TGTP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------
 * The reference reports invalid input by throwing std::runtime_error /
 * std::invalid_argument at load time (graph.h:34,55,69,74; aesc.h:53-61,
 * 87-89; main.cpp:27-30).  Across this ABI the same conditions return a
 * negative code and leave a message in tgtp_ctx_last_error(). */
#define TGTP_OK 0
#define TGTP_E_ARG (-1)         /* invalid argument / input (runtime_error, invalid_argument) */
#define TGTP_E_HIP (-2)         /* HIP runtime error */
#define TGTP_E_NOMEM (-3)       /* host or device allocation failed (bad_alloc) */
#define TGTP_E_IO (-4)          /* file could not be opened / parsed */
#define TGTP_E_UNSUPPORTED (-5) /* shape outside what this build's kernels handle */

typedef struct tgtp_ctx tgtp_ctx;       /* owns a HIP stream, events, scratch */
typedef struct tgtp_dgraph tgtp_dgraph; /* device-resident CSR (HBM) */

/* Host CSR view (the fields of fuser::graph_t, graph.h:13-15).  Rows sorted,
 * ids < n.  The walk entry points take any such CSR; tgtp_taus, the
 * tgtp_score family and the batch also need both directions of every edge,
 * as graph.h's loader stores them (TGTP_E_ARG otherwise). */
typedef struct tgtp_graph {
  uint64_t n;          /* vertices */
  uint64_t m;          /* CSR entries (both directions) */
  const uint32_t* idx; /* [n+1] */
  const uint32_t* adj; /* [m]   */
} tgtp_graph;

/* Eigenpairs (EigenVecs<float>, aesc.h:43-81). */
typedef struct tgtp_eigens {
  int32_t omega;          /* number of pairs, >= 3 */
  const double* values;   /* [omega] */
  const float* vectors;   /* [n*omega] vertex-major */
} tgtp_eigens;

/* Arithmetic of tgtp_score.
 *  FAITHFUL  : every step in the reference's own float/double arithmetic and
 *              order; scores equal the CPU restatement bit for bit.
 *  FAST_FP32 : tau projection as a batched contraction with FP32 operands and
 *              FP32 accumulation; per-edge score sums in FP64.
 *  FAST_BF16 : as FAST_FP32 with BF16 eigenvector operands. */
enum {
  TGTP_MODE_FAITHFUL = 0,
  TGTP_MODE_FAST_FP32 = 1,
  TGTP_MODE_FAST_BF16 = 2
};

/* Tau refinement rule of the FAST modes (a flag OR'ed into the mode).
 * The reference refines tau with get_tau(..., delta, upsilon), where delta is
 * the top-omega tail of the series after t steps, and when delta >= eps/2 it
 * bounds the remaining terms against eps/2 instead of eps/2 - delta
 * (aesc.h:130): the tail is then left out of the bound and tau can drop to 1
 * on an edge whose tail alone exceeds eps/2 (bridges into sparse parts of a
 * graph; errors up to 0.5 on the batched workload's graphs).  By default the
 * FAST modes skip such a refinement step (t advances, tau is kept), so every
 * tau satisfies  delta(tau) + rest(tau) <= eps/2.  TGTP_TAU_REFERENCE selects
 * the reference's rule in FAST modes; FAITHFUL always uses it. */
#define TGTP_TAU_REFERENCE 0x100

/* The 2-way walks of the reference's baseline tgtp (AESC::tgtp of
 * aesc_baseline.h, the comparator driven by benchmark_baseline.cpp), a flag
 * OR'ed into TGTP_MODE_FAITHFUL: instead of the SABA bouquets, each walk
 * draws k = rndgen() % degree for the walker from s, then for the walker
 * from v, at every hop, from one std::mt19937 seeded 42 at the start of the
 * call and consumed in source and position order (aesc_baseline.h:17, 261,
 * 363-379); each position's differences are summed in one float chain.
 * Scores equal the reference baseline's bit for bit.  FAST modes refuse it. */
#define TGTP_WALK2_MT19937 0x200

/* AESCConfig (aesc.h:98-102); the reference CLI sets delta = 1/n
 * (main.cpp:36).  gamma >= 1, 0 < epsilon < 1. */
typedef struct tgtp_params {
  double epsilon;
  double delta;
  int32_t gamma;
  int32_t mode; /* TGTP_MODE_* */
} tgtp_params;

/* Timings of the last call, from HIP events on the context's stream. */
typedef struct tgtp_stats {
  double kernel_ms;      /* device time of the call's kernels (event to event) */
  double e2e_ms;         /* host wall time including H2D / D2H copies */
  double walk_ms;        /* device time of the walk kernel(s) alone */
  double taus_ms;        /* tgtp: tau stage */
  double diffuse_ms;     /* tgtp: per-source diffusion + walk-count stage */
  uint64_t transitions;  /* walks: n_sources_with_degree * n_walks * depth; tgtp: walker-hops */
  uint64_t work_items;   /* tgtp: pushed edges (diffusion), walks: bouquets */
  uint64_t host_fixups;  /* tgtp: results resolved on the host (documented cases) */
} tgtp_stats;

/* ---- context ------------------------------------------------------------ */
int tgtp_ctx_create(int device, tgtp_ctx** out);
void tgtp_ctx_destroy(tgtp_ctx* ctx);
const char* tgtp_ctx_last_error(const tgtp_ctx* ctx);
const char* tgtp_strerror(int code);

/* ---- device-resident graph (inputs stay in HBM across calls) ------------- */
int tgtp_graph_upload(tgtp_ctx* ctx, const tgtp_graph* g, tgtp_dgraph** out);
void tgtp_graph_free(tgtp_ctx* ctx, tgtp_dgraph* dg);

/* ---- SABA 1-way walks ------------------------------------------------------
 * For every source s in [src_begin, src_end) with degree > 0, n_walks walks
 * of exactly `depth` hops on the ordered uniform 32-bit grid, grouped into
 * 8-lane bouquets, hashed steps with a whole-bouquet SplitMix32 fallback
 * (rw_experiment_saba.cpp:65-145).  Counters are uint32 per vertex (the
 * reference keeps per-thread float counters, rw_experiment_saba.cpp:192).
 * counters_out [n] may be NULL; visit_count = sum(count), checksum =
 * sum((v+1)*count[v]) in uint64 (rw_experiment_saba.cpp:203-206).
 * n_walks must be a multiple of 8 (rw_experiment_saba.cpp:160-163); 0 walks
 * nothing and returns zero counters and sums, as the reference does (:186).
 * stats->kernel_ms covers counter clearing + walks + reduction, the window
 * of t_rw_s (rw_experiment_saba.cpp:184-211) minus host work. */
int tgtp_saba_walks(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin,
                    uint32_t src_end, uint32_t depth, uint32_t n_walks,
                    uint32_t* counters_out, uint64_t* visit_count,
                    uint64_t* checksum, tgtp_stats* stats);

/* ---- MT19937 comparator walks ----------------------------------------------
 * The reference's baselines of the walk comparison, on the GPU:
 *   TGTP_WALK_MT_REGULAR <- scalar_randomwalk + driver   src/rw_experiment.cpp:34-47, 87-108
 *   TGTP_WALK_MT_SIMD    <- simd_randomwalk + driver     src/rw_experiment_simd.cpp:44-82
 * Every source s in [src_begin, src_end) with degree > 0 draws from its own
 * std::mt19937 seeded with 42 + s; n_walks walks of exactly `depth` hops,
 * next vertex adj[row + draw % degree]; regular: walk i, hop h takes draw
 * i*depth + h; simd: bouquets of 8 walks, hop h of bouquet b takes draws
 * b*8*depth + 8h + lane.  Counters and sums as tgtp_saba_walks.  simd needs
 * n_walks % 8 == 0 (TGTP_E_ARG otherwise); depth up to 3472 (regular) / 434
 * (simd), TGTP_E_UNSUPPORTED beyond.  In a non-symmetric CSR a walker that
 * reaches a vertex without neighbours stops there (the reference divides by
 * zero); the draw indices of all other walks are unchanged. */
enum { TGTP_WALK_MT_REGULAR = 0, TGTP_WALK_MT_SIMD = 1 };
int tgtp_mt_walks(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin,
                  uint32_t src_end, uint32_t depth, uint32_t n_walks, int32_t variant,
                  uint32_t* counters_out, uint64_t* visit_count, uint64_t* checksum,
                  tgtp_stats* stats);

/* tgtp_mt_walks without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, checks and results as tgtp_mt_walks; the device counters and
 * sums live in a static 16 MiB scratch array that each GPU holds until the
 * process ends, so graphs of at most 4,194,304 vertices are accepted
 * (TGTP_E_UNSUPPORTED beyond).  Every call on one GPU uses that GPU's array
 * (cleared at the start of each call): calls that overlap on one GPU, from
 * any context or thread, race. */
int tgtp_mt_walks_noalloc(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin,
                          uint32_t src_end, uint32_t depth, uint32_t n_walks, int32_t variant,
                          uint32_t* counters_out, uint64_t* visit_count, uint64_t* checksum,
                          tgtp_stats* stats);

/* ---- tau bounds (calculate_taus, aesc.h:164-216) ---------------------------
 * eps_half is float(epsilon/2) as aesc.h:122 passes it.  taus_out [m]. */
int tgtp_taus(tgtp_ctx* ctx, const tgtp_graph* g, const tgtp_eigens* e,
              float eps_half, int32_t mode, int32_t* taus_out,
              tgtp_stats* stats);

/* ---- full TGT+ (AESC::tgtp, aesc.h:120-127) --------------------------------
 * scores_out [m] float32 in CSR order: exactly the scores.bin payload
 * (main.cpp:41).  Inputs are host arrays; they are copied to HBM. */
int tgtp_score(tgtp_ctx* ctx, const tgtp_graph* g, const tgtp_eigens* e,
               const tgtp_params* p, float* scores_out, tgtp_stats* stats);

/* Same, on a graph already resident in HBM (eigenpairs are uploaded per call
 * unless the same vectors and values pointers and omega were passed on the
 * previous call with this dgraph; contents are not compared, so eigenpairs
 * changed in place need a fresh upload of the dgraph or other buffers). */
int tgtp_score_dev(tgtp_ctx* ctx, const tgtp_dgraph* dg, const tgtp_eigens* e,
                   const tgtp_params* p, float* scores_out, tgtp_stats* stats);

/* ---- batched form -------------------------------------------------------------
 * No counterpart in the reference, which scores one graph per process
 * (benchmark.cpp:29-50).  Scores `count` graphs of uneven size together:
 * in FAST modes they run as one disjoint-union CSR -- one tau projection
 * launch, one persistent per-source launch over a work list spanning all
 * graphs (largest first), one symmetrization -- optionally autotuned and
 * captured into a HIP graph that every tgtp_batch_run replays.  All graphs
 * share one mode; FAST batches need one omega in [3, 128].  FAITHFUL batches
 * run the graphs one after another (bit-exact as tgtp_score).  Per-graph
 * scores equal what tgtp_score_dev returns for that graph with the same
 * tuning knobs.  scores_out[k] holds graphs[k].m floats.  FAST batches copy
 * every input at create; FAITHFUL batches copy the graphs but read the
 * eigenpair arrays when they run, so those must outlive the batch. */
typedef struct tgtp_batch tgtp_batch;
#define TGTP_BATCH_AUTOTUNE 1 /* time candidate launch shapes once, keep the best */
#define TGTP_BATCH_GRAPH 2    /* capture the run into a hipGraph, replay it */
int tgtp_batch_create(tgtp_ctx* ctx, int32_t count, const tgtp_graph* graphs,
                      const tgtp_eigens* eigens, const tgtp_params* params, int32_t flags,
                      tgtp_batch** out);
int tgtp_batch_run(tgtp_ctx* ctx, tgtp_batch* batch, float* const* scores_out,
                   tgtp_stats* stats);
void tgtp_batch_destroy(tgtp_ctx* ctx, tgtp_batch* batch);
/* Launch shape the batch runs with (after autotuning, if requested). */
int tgtp_batch_config(const tgtp_batch* batch, uint32_t* light_degree, uint32_t* small_bouquets,
                      uint32_t* slots, uint32_t* max_slots);
/* Diagnostic run of a FAST batch (uncaptured, its tuned launch shape) with
 * an event pair around every launch of the diffusion round kernel, the stop
 * kernel and the post kernel, on the stream each runs on.  round_bytes are
 * the algorithmic bytes of all round launches: per source and round after
 * the first, 4 bytes per edge of its graph (its value gathered once per
 * neighbour entry) and 4 per vertex (its new value written); per block of 64
 * sources and round, 4 bytes per edge (neighbour ids) and 8 per vertex (row
 * descriptors).  No counterpart in the reference. */
typedef struct tgtp_profile {
  double total_ms;          /* the whole run, event to event */
  double taus_ms;           /* tau projection */
  double round_ms;          /* sum over round-kernel launches */
  uint64_t round_launches;
  uint64_t round_bytes;
  double stop_ms;
  uint64_t stop_launches;
  double post_ms;
  uint64_t post_launches;
  uint64_t source_rounds;   /* sum over sources of the rounds each took */
} tgtp_profile;
int tgtp_batch_profile(tgtp_ctx* ctx, tgtp_batch* batch, tgtp_profile* out);
/* One-shot: create (no tuning, no capture), run, destroy. */
int tgtp_score_batch(tgtp_ctx* ctx, int32_t count, const tgtp_graph* graphs,
                     const tgtp_eigens* eigens, const tgtp_params* params,
                     float* const* scores_out, tgtp_stats* stats);

/* ---- source subsets ----------------------------------------------------------
 * No counterpart in the reference, which scores every source; it serves
 * sampled measurements of graphs whose full run takes minutes (the
 * reference's per-source work is independent, aesc.h:456-527, so a uniform
 * sample of sources extrapolates).  tgtp_score_dev restricted to the sources
 * of n_ranges half-open ranges ranges[2k] .. ranges[2k+1] (non-empty,
 * ascending, disjoint, within [0, n)): taus for every edge, the per-source
 * pipeline for the listed sources only.  rows_out [m] (nullable) receives
 * the row value of every CSR position whose source is listed (preds before
 * symmetrization, aesc.h:532: FAITHFUL the float row, bit for bit the full
 * run's; FAST the FP64 row sum rounded to float); scores_out [m] (nullable)
 * the score of every position whose source and target are both listed,
 * equal to what the full run returns there.  All other entries are left
 * unchanged.  stats->work_items = sources run. */
int tgtp_score_sources(tgtp_ctx* ctx, const tgtp_dgraph* dg, const tgtp_eigens* e,
                       const tgtp_params* p, uint32_t n_ranges, const uint32_t* ranges,
                       float* rows_out, float* scores_out, tgtp_stats* stats);

/* Diagnostic form of tgtp_score_dev that also returns the per-stage values
 * of the pipeline (each output pointer may be NULL): taus [m] (calculate_taus,
 * aesc.h:164-216), nr [m] walk count per position (0 where tau <= ell,
 * aesc.h:519-521), walk_sum [m] raw 2-way walk sum (aesc.h:523),
 * preds_row [m] row predictions before symmetrization (aesc.h:532), and per
 * source ell / edge_max / global max / global min / final frontier size
 * [n] (aesc.h:460-506).  No counterpart in the reference: it exposes the
 * intermediate values of AESC::tgtp for stage-by-stage parity checks. */
int tgtp_score_trace(tgtp_ctx* ctx, const tgtp_dgraph* dg, const tgtp_eigens* e,
                     const tgtp_params* p, float* scores_out, tgtp_stats* stats,
                     int32_t* taus, uint64_t* nr, float* walk_sum, float* preds_row,
                     int32_t* ell, float* edge_max, float* gmax, float* gmin,
                     uint32_t* q_size);

#ifdef __cplusplus
}
#endif
#endif /* TGTP_HIP_H */
